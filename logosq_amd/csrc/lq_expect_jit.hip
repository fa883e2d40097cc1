// Pauli-sum expectations from 128 KiB LDS tiles with kernels specialised at
// run time: the fast path of PauliObservable::expectation
// (src/optimization/observable.rs:95-147, 228-231) for large states.
//
// A layout is a tile geometry: the L lowest index bits (L = 3..5, so every
// global access covers whole 128-byte lines) plus K - L chosen high bits,
// K = 13 (128 KiB of complex128 / 64 KiB of complex64 amplitudes), one
// 1024-thread block per CU.  The layouts cover the sum's flip groups (terms
// sharing their X/Y bits), greedily, choosing L per layout; the diagonal
// group (I/Z terms) rides along in the first layout, so a local Hamiltonian
// costs a few reads of the state in all (the 30-qubit XYZ ring: three).
//
// Inside a layout every thread holds A = 8 amplitudes of a "window": the subcube over w = log2(A) tile bits, its
// other tile bits given by the thread index.  A flip group whose X/Y bits lie
// in the window pairs amplitudes held in registers, and its terms' signs
// within the window are compile-time constants; signs from bits outside the
// window are per thread (fixed for the kernel) or per tile (uniform, scalar
// unit).  A window pass gathers the thread's amplitudes from LDS once for all
// its groups.  Groups that fit the tile but no window read their pairs from
// LDS one by one.  Each layout's kernel is written as text from the sum's
// structure (not its coefficients, which come from a table, so one kernel
// serves every coupling and every instance of a batch) and compiled with
// hipRTC under the library's run-time specialisation policy.
//
// For a pair (i, j = i ^ x) with bit h of i clear and a term t with sign
// s_t(i) = (-1)^{popc(i & zy_t)}, both rows together contribute
// s_t(i) * 2 cr_t * Re(conj(psi_i) psi_j) if popc(x & zy_t) is even and
// s_t(i) * (-2 ci_t) * Im(conj(psi_i) psi_j) if odd; a diagonal term
// contributes s_t(i) * cr_t * |psi_i|^2.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lq_expect_jit.h"
#include "lq_fused.h"

namespace lq {

namespace {

constexpr int kXMaxTerms = 384;   // terms per layout kernel (larger sums keep the generic kernels)

}  // namespace

struct XKernel {
    int seen = 0;
    bool failed = false;
    std::vector<char> code;
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
};

namespace {

std::unordered_map<std::string, std::unique_ptr<XKernel>> g_xk;
constexpr size_t kXMaxKernels = 1024;
int g_cu_count = 0;

// 2^13 amplitudes per tile for both types (128 KiB complex128, 64 KiB
// complex64): with 8 amplitudes per thread a block of 1024 threads stays
// within 128 VGPRs (complex64 amplitudes are widened to f64 for exact
// products, which at 16 per thread spilled).
int tile_max_bits(int) { return 13; }
// log2 of the amplitudes each thread holds: the window size.  8 by default,
// so a tile takes 1024 threads; LQ_EXPECT_AMPS_LOG=4 gives 16 (512 threads).
int window_bits(int dtype) {
    static int over = [] {
        const char* e = std::getenv("LQ_EXPECT_AMPS_LOG");
        return e ? std::atoi(e) : 0;
    }();
    (void)dtype;
    return over == 4 ? 4 : 3;
}
// fewest contiguous low bits: a 128-byte line
int low_min(int dtype) { return dtype == LQ_C128 ? 3 : 4; }
// bits of one LDS swizzle digit: 128 bytes of amplitudes
int digit_bits(int dtype) { return dtype == LQ_C128 ? 3 : 4; }

uint32_t swz_host(uint32_t e, int K, int dtype) {
    const int d = digit_bits(dtype);
    uint32_t f = 0;
    for (int s = d; s < K; s += d) f ^= e >> s;
    return e ^ (f & ((1u << d) - 1u));
}

}  // namespace


namespace {

// ---- planning ----------------------------------------------------------
// Tile bits for one layout: for each L from 5 down to the line minimum, take
// groups greedily by the fewest new high bits (low_first: by the lowest top
// high bit, then the fewest new bits); keep the L covering the most groups
// (ties: the larger L).
void choose_tile(int n, int dtype, const PauliSum& ps, const std::vector<int>& todo, bool low_first, int& L_out,
                 uint64_t& H_out) {
    const int K = std::min(tile_max_bits(dtype), n);
    int best = -1;
    for (int L = std::min(5, K); L >= std::min(low_min(dtype), K); --L) {
        const int HB = K - L;
        const uint64_t low = (1ull << L) - 1ull;
        uint64_t H = 0;
        std::vector<char> taken(todo.size(), 0);
        for (;;) {
            int pick = -1, pc = 65, ptop = 65;
            for (size_t i = 0; i < todo.size(); ++i) {
                if (taken[i]) continue;
                const uint64_t need = ps.groups[todo[i]].x & ~low & ~H;
                const int c = __builtin_popcountll(need);
                const uint64_t HH = H | need;
                const int top = HH ? 63 - __builtin_clzll(HH) : 0;
                if (__builtin_popcountll(HH) > HB) continue;
                if (low_first ? (top < ptop || (top == ptop && c < pc)) : c < pc) {
                    pick = (int)i;
                    pc = c;
                    ptop = top;
                }
            }
            if (pick < 0) break;
            H |= ps.groups[todo[pick]].x & ~low;
            taken[pick] = 1;
        }
        for (int b = L; b < n && __builtin_popcountll(H) < HB; ++b) H |= 1ull << b;
        int cov = 0;
        for (int g : todo) cov += (ps.groups[g].x & ~(low | H)) == 0;
        if (cov > best) {
            best = cov;
            L_out = L;
            H_out = H;
        }
    }
}

// Cover a layout's groups (local flip masks) with windows of w bits; groups
// wider than w go to `pairs`.  Returns window bit sets and, per group, its
// window (-1: pairs).
void choose_windows(int K, int w, const std::vector<uint32_t>& xl, std::vector<std::vector<int>>& wins,
                    std::vector<int>& where) {
    where.assign(xl.size(), -1);
    std::vector<int> left;
    for (size_t i = 0; i < xl.size(); ++i)
        if (__builtin_popcount(xl[i]) <= w) left.push_back((int)i);
    while (!left.empty()) {
        uint32_t bestW = 0;
        int bestc = -1;
        for (int s : left) {
            uint32_t W = xl[s];
            while (__builtin_popcount(W) < w) {
                int pb = -1, pg = -1;
                for (int b = 0; b < K; ++b) {
                    if (W & (1u << b)) continue;
                    int g = 0;
                    for (int i : left) g += (xl[i] & ~(W | (1u << b))) == 0;
                    if (g > pg) {
                        pg = g;
                        pb = b;
                    }
                }
                W |= 1u << pb;
            }
            int c = 0;
            for (int i : left) c += (xl[i] & ~W) == 0;
            if (c > bestc) {
                bestc = c;
                bestW = W;
            }
        }
        std::vector<int> bits;
        for (int b = 0; b < K; ++b)
            if (bestW & (1u << b)) bits.push_back(b);
        const int id = (int)wins.size();
        wins.push_back(bits);
        std::vector<int> still;
        for (int i : left) {
            if ((xl[i] & ~bestW) == 0) where[i] = id;
            else still.push_back(i);
        }
        left.swap(still);
    }
}

// Thread index bits -> the tile bits outside the window.  The lowest
// digit_bits of them go to bits of distinct residues modulo the swizzle
// digit (so the eight (complex128) / sixteen (complex64) lanes that read one
// 128-byte LDS row hit distinct slots), the rest ascending.
std::vector<int> thread_bits(int K, int dtype, const std::vector<int>& wbits) {
    const int d = digit_bits(dtype);
    std::vector<int> free;
    for (int b = 0; b < K; ++b)
        if (std::find(wbits.begin(), wbits.end(), b) == wbits.end()) free.push_back(b);
    std::vector<int> first;
    uint32_t res = 0;
    for (int b : free)
        if ((int)first.size() < d && !(res & (1u << (b % d)))) {
            first.push_back(b);
            res |= 1u << (b % d);
        }
    std::vector<int> out = first;
    for (int b : free)
        if (std::find(first.begin(), first.end(), b) == first.end()) out.push_back(b);
    return out;
}

}  // namespace

bool expect_jit_wanted(int n, int dtype) {
    if (jit_mode() == LQ_JIT_OFF) return false;
    if (n < window_bits(dtype) + 6) return false;  // at least one wavefront per tile
    return expect_tiles_wanted(n, dtype);
}

bool expect_jit_wanted_batch(int n, int dtype, uint64_t ninst) {
    if (jit_mode() == LQ_JIT_OFF) return false;
    if (n < window_bits(dtype) + 6) return false;
    return expect_tiles_batch_wanted(n, dtype, ninst);
}

// Plan the layouts of a compiled sum on an n-qubit state with one greedy
// criterion (choose_tile).  False when the sum does not suit the specialised
// kernels (no flip group fits a layout, or too many terms).
static bool plan_expect_jit_with(int n, int dtype, const PauliSum& ps, bool low_first, XPlan& xp) {
    xp.n = n;
    xp.dtype = dtype;
    xp.layouts.clear();
    xp.rest.clear();
    xp.diag_in = false;
    const int K = std::min(tile_max_bits(dtype), n);
    const int w = window_bits(dtype);
    std::vector<int> todo;
    int diag = -1;
    for (int g = 0; g < (int)ps.groups.size(); ++g) {
        if (ps.groups[g].x == 0) diag = g;
        else todo.push_back(g);
    }
    while (!todo.empty()) {
        int L = 0;
        uint64_t H = 0;
        choose_tile(n, dtype, ps, todo, low_first, L, H);
        const uint64_t low = (1ull << L) - 1ull, tilebits = low | H;
        std::vector<int> in, out;
        for (int g : todo) ((ps.groups[g].x & ~tilebits) == 0 ? in : out).push_back(g);
        if (in.empty()) break;
        XLayout lay;
        lay.K = K;
        lay.L = L;
        for (int b = L; b < n; ++b)
            if (H & (1ull << b)) lay.hpos.push_back(b);
        auto local = [&](uint64_t m) -> uint32_t {
            uint32_t r = (uint32_t)(m & low);
            for (size_t j = 0; j < lay.hpos.size(); ++j)
                if (m & (1ull << lay.hpos[j])) r |= 1u << (L + (int)j);
            return r;
        };
        std::vector<uint32_t> xl;
        for (int g : in) xl.push_back(local(ps.groups[g].x));
        std::vector<std::vector<int>> wins;
        std::vector<int> where;
        choose_windows(K, w, xl, wins, where);
        const bool with_diag = xp.layouts.empty() && diag >= 0;
        if (with_diag && wins.empty()) {
            std::vector<int> b;
            for (int i = 0; i < w; ++i) b.push_back(i);
            wins.push_back(b);
        }
        for (auto& wb : wins) {
            XWindow W;
            W.bits = wb;
            W.tbits = thread_bits(K, dtype, wb);
            lay.windows.push_back(W);
        }
        auto make_term = [&](const PauliTermDev& pt, uint64_t x, const XWindow* W) {
            XTerm T;
            T.gi = (int)lay.coef.size();
            T.zyo = pt.zy & ~tilebits;
            const uint32_t zl = local(pt.zy & tilebits);
            T.zw = 0;
            T.zyr = zl;
            if (W) {
                for (size_t k = 0; k < W->bits.size(); ++k)
                    if (zl & (1u << W->bits[k])) T.zw |= 1u << k;
                uint32_t wm = 0;
                for (int b : W->bits) wm |= 1u << b;
                T.zyr = zl & ~wm;
            }
            if (x == 0) {
                T.im = 0;
                lay.coef.push_back(pt.cr);
            } else {
                T.im = __builtin_popcountll(x & pt.zy) & 1;
                lay.coef.push_back(T.im ? -2.0 * pt.ci : 2.0 * pt.cr);
            }
            return T;
        };
        for (size_t i = 0; i < in.size(); ++i) {
            const PauliGroup& grp = ps.groups[in[i]];
            XGroup G;
            G.xl = xl[i];
            if (where[i] >= 0) {
                XWindow& W = lay.windows[where[i]];
                G.xw = 0;
                for (size_t k = 0; k < W.bits.size(); ++k)
                    if (G.xl & (1u << W.bits[k])) G.xw |= 1u << k;
                G.hw = 31 - __builtin_clz(G.xw);
                for (int t = grp.first; t < grp.first + grp.count; ++t) G.terms.push_back(make_term(ps.terms[t], grp.x, &W));
                W.groups.push_back(G);
            } else {
                G.xw = 0;
                G.hw = 31 - __builtin_clz(G.xl);
                for (int t = grp.first; t < grp.first + grp.count; ++t)
                    G.terms.push_back(make_term(ps.terms[t], grp.x, nullptr));
                lay.pairs.push_back(G);
            }
        }
        if (with_diag) {
            const PauliGroup& grp = ps.groups[diag];
            for (int t = grp.first; t < grp.first + grp.count; ++t)
                lay.windows[0].diag.push_back(make_term(ps.terms[t], 0, &lay.windows[0]));
            xp.diag_in = true;
        }
        if (lay.coef.size() > (size_t)kXMaxTerms) return false;
        xp.layouts.push_back(std::move(lay));
        todo.swap(out);
    }
    xp.rest = todo;
    return !xp.layouts.empty();
}

// Reads of the state a plan makes (its layouts, then the groups left to orbit
// items, four per read) and the pages its tiles span: a tile whose high bits
// reach above a 2 MiB page touches 2^(bits above) pages, and a layout whose
// tiles span many pages runs at the rate of its address translations (30
// qubits, the ring's middle layout: bits 20-29 in the tile, 1024 pages, 4.3 ms
// against 2.9 ms for bits 12-21, 32 pages).
static std::pair<size_t, double> plan_cost(const XPlan& xp) {
    const int page_bits = xp.dtype == LQ_C128 ? 17 : 18;
    double pages = 0.0;
    for (const XLayout& l : xp.layouts) {
        int above = 0;
        for (int b : l.hpos) above += b >= page_bits;
        pages += std::ldexp(1.0, above);
    }
    return {xp.layouts.size() + (xp.rest.size() + 3) / 4, pages};
}

// Plan with both greedy criteria; keep the plan with fewer reads, and at
// equal reads the low-first plan only when its tiles span clearly fewer pages
// (a quarter fewer: near-equal spans measured no better, complex64 ring 7.3
// against 7.6 ms).
bool plan_expect_jit(int n, int dtype, const PauliSum& ps, XPlan& xp) {
    XPlan lo;
    const bool a = plan_expect_jit_with(n, dtype, ps, false, xp);
    const bool b = plan_expect_jit_with(n, dtype, ps, true, lo);
    if (!b) return a;
    if (!a) {
        xp = std::move(lo);
        return true;
    }
    const auto ca = plan_cost(xp), cb = plan_cost(lo);
    if (cb.first < ca.first || (cb.first == ca.first && cb.second <= 0.75 * ca.second)) xp = std::move(lo);
    return true;
}

namespace {

// ---- kernel text ---------------------------------------------------------
std::string hex(uint64_t v) {
    char b[32];
    std::snprintf(b, sizeof(b), "0x%llxull", (unsigned long long)v);
    return b;
}

// Kernel text of one layout.  Coefficient table entries used by the text:
// the layout's raw coefficients (lay.coef), then `derived` sums of them with
// signs (filled here), so that every weight that does not depend on the tile
// or the thread is one table entry, a scalar operand of the FMA.
std::string layout_source(const XLayout& lay, int n, int dtype, std::vector<Combo>* derived_out) {
    const int K = lay.K, L = lay.L;
    const int w = window_bits(dtype), A = 1 << w;
    const int NTL = K - w, NT = 1 << NTL;
    const uint64_t ntiles = 1ull << (n - K);
    const char* V = dtype == LQ_C128 ? "double2" : "float2";
    const char* S = dtype == LQ_C128 ? "double" : "float";
    const int nraw = (int)lay.coef.size();
    std::vector<Combo> derived;
    std::map<Combo, int> dmap;
    // table entry holding sum_k sign_k * coef[gi_k] (combo sorted); -1 for an empty sum
    auto entry = [&](Combo c) -> int {
        std::sort(c.begin(), c.end());
        if (c.empty()) return -1;
        if (c.size() == 1 && c[0].second == 1) return c[0].first;
        auto it = dmap.find(c);
        if (it != dmap.end()) return it->second;
        const int id = nraw + (int)derived.size();
        derived.push_back(c);
        dmap[c] = id;
        return id;
    };
    std::vector<int> gpos(K);  // global position of each local bit
    for (int b = 0; b < K; ++b) gpos[b] = b < L ? b : lay.hpos[b - L];
    auto spread_local = [&](uint32_t e) {
        uint64_t g = 0;
        for (int b = 0; b < K; ++b)
            if (e & (1u << b)) g |= 1ull << gpos[b];
        return g;
    };
    auto swz = [&](uint32_t e) { return swz_host(e, K, dtype); };
    const bool c64 = dtype == LQ_C64;
    auto cx = [&](const std::string& e) { return c64 ? "(double)" + e : e; };
    auto G = [](int i) { return "G[" + std::to_string(i) + "]"; };
    auto tsign = [&](uint64_t zyo) { return "(bool)(__popcll(base & " + hex(zyo) + ") & 1)"; };
    int tmp = 0;
    std::ostringstream pre, body;
    for (size_t wi = 0; wi < lay.windows.size(); ++wi) {
        const XWindow& W = lay.windows[wi];
        const std::string tb = "tb" + std::to_string(wi), sb = "sb" + std::to_string(wi);
        pre << "  unsigned " << tb << " = 0;\n";
        for (int i = 0; i < NTL; ++i) pre << "  " << tb << " |= ((tid >> " << i << ") & 1u) << " << W.tbits[i] << ";\n";
        pre << "  const unsigned " << sb << " = swz(" << tb << ");\n";
        auto thr_sign = [&](uint32_t zyr) { return "(__popc(" + tb + " & " + std::to_string(zyr) + "u) & 1)"; };
        // A term's signed coefficient for the row at window offset o is
        // chi(o) * [thread sign] * [tile sign] * coef.  Returns, for a set of
        // terms with their chi(o) signs, the expression of the sum: one table
        // entry for the sign-free terms, a hoisted per-thread sum for the
        // thread-signed ones, and per-tile terms.
        std::map<std::string, std::string> hoisted, per_tile;
        std::ostringstream tilec, work;
        auto weight = [&](const std::vector<std::pair<const XTerm*, int>>& ts) -> std::string {
            Combo cst;
            std::vector<std::string> thr, til;
            for (auto& [T, sg] : ts) {
                if (!T->zyo && !T->zyr) {
                    cst.emplace_back(T->gi, sg);
                } else if (!T->zyo) {
                    thr.push_back("(" + thr_sign(T->zyr) + " ? -" + G(T->gi) + " : " + G(T->gi) + ")" +
                                  (sg < 0 ? " * -1.0" : ""));
                } else {
                    std::string cond = tsign(T->zyo);
                    if (T->zyr) {
                        const std::string h = "h" + std::to_string(tmp++);
                        pre << "  const bool " << h << " = " << thr_sign(T->zyr) << ";\n";
                        cond = "(" + h + " ^ " + cond + ")";
                    }
                    til.push_back("(" + cond + " ? " + (sg < 0 ? "" : "-") + G(T->gi) + " : " + (sg < 0 ? "-" : "") +
                                  G(T->gi) + ")");
                }
            }
            std::string e;
            const int ce = entry(cst);
            if (ce >= 0) e = G(ce);
            if (!thr.empty()) {
                std::string key;
                for (auto& x : thr) key += (key.empty() ? "" : " + ") + x;
                auto it = hoisted.find(key);
                if (it == hoisted.end()) {
                    const std::string nm = "q" + std::to_string(tmp++);
                    pre << "  const double " << nm << " = " << key << ";\n";
                    it = hoisted.emplace(key, nm).first;
                }
                e += (e.empty() ? "" : " + ") + it->second;
            }
            if (!til.empty()) {
                std::string key = e;
                for (auto& x : til) key += (key.empty() ? "" : " + ") + x;
                auto it = per_tile.find(key);
                if (it == per_tile.end()) {
                    const std::string nm = "u" + std::to_string(tmp++);
                    tilec << "    const double " << nm << " = " << key << ";\n";
                    it = per_tile.emplace(key, nm).first;
                }
                return it->second;
            }
            return e;
        };
        work << "    " << V << " v[" << A << "];\n";
        for (int o = 0; o < A; ++o) {
            uint32_t off = 0;
            for (int k = 0; k < w; ++k)
                if (o & (1 << k)) off |= 1u << W.bits[k];
            work << "    v[" << o << "] = tile[" << sb << " ^ " << swz(off) << "u];\n";
        }
        for (const XGroup& Gp : W.groups) {
            for (int o = 0; o < A; ++o) {
                if (o & (1 << Gp.hw)) continue;
                const int p = o ^ (int)Gp.xw;
                const std::string a = "v[" + std::to_string(o) + "]", b = "v[" + std::to_string(p) + "]";
                for (int im = 0; im < 2; ++im) {
                    std::vector<std::pair<const XTerm*, int>> ts;
                    for (const XTerm& T : Gp.terms)
                        if (T.im == im) ts.emplace_back(&T, (__builtin_popcount((uint32_t)o & T.zw) & 1) ? -1 : 1);
                    if (ts.empty()) continue;
                    const std::string wgt = weight(ts);
                    if (wgt.empty()) continue;
                    if (!im)
                        work << "    acc = __builtin_fma(" << cx(a + ".x") << " * " << cx(b + ".x") << " + " << cx(a + ".y")
                             << " * " << cx(b + ".y") << ", " << wgt << ", acc);\n";
                    else
                        work << "    acc = __builtin_fma(" << cx(a + ".x") << " * " << cx(b + ".y") << " - " << cx(a + ".y")
                             << " * " << cx(b + ".x") << ", " << wgt << ", acc);\n";
                }
            }
        }
        if (!W.diag.empty()) {
            // per window sign pattern zw: P_zw = sum_o (-1)^{popc(o & zw)} |v_o|^2, accumulated as
            // each |v_o|^2 is formed, then one FMA with the class's summed coefficient
            std::map<uint32_t, std::vector<std::pair<const XTerm*, int>>> cls;
            for (const XTerm& T : W.diag) cls[T.zw].emplace_back(&T, 1);
            std::vector<std::string> pn;
            for (auto& kv : cls) {
                pn.push_back("P" + std::to_string(tmp++));
                work << "    double " << pn.back() << " = 0.0;\n";
            }
            for (int o = 0; o < A; ++o) {
                const std::string a = "v[" + std::to_string(o) + "]";
                work << "    { const double p = " << cx(a + ".x") << " * " << cx(a + ".x") << " + " << cx(a + ".y")
                     << " * " << cx(a + ".y") << ";";
                size_t ci = 0;
                for (auto& kv : cls) {
                    const bool neg = __builtin_popcount((uint32_t)o & kv.first) & 1;
                    work << " " << pn[ci++] << (neg ? " -= p;" : " += p;");
                }
                work << " }\n";
            }
            size_t ci = 0;
            for (auto& [zw, ts] : cls) work << "    acc = __builtin_fma(" << pn[ci++] << ", " << weight(ts) << ", acc);\n";
        }
        body << "    {\n" << tilec.str() << work.str() << "    }\n";
    }
    // groups in the tile but in no window: pairs from LDS, in-tile signs per pair
    for (const XGroup& Gp : lay.pairs) {
        std::ostringstream tilec, work;
        work << "    for (int u = 0; u < " << (A / 2) << "; ++u) {\n";
        work << "      const unsigned k = tid + (unsigned)u * " << NT << "u;\n";
        work << "      const unsigned e = ((k >> " << Gp.hw << ") << " << (Gp.hw + 1) << ") | (k & " << ((1u << Gp.hw) - 1u)
             << "u);\n";
        work << "      const unsigned se = swz(e);\n";
        work << "      const " << V << " a = tile[se], b = tile[se ^ " << swz(Gp.xl) << "u];\n";
        work << "      const double zr = " << cx("a.x") << " * " << cx("b.x") << " + " << cx("a.y") << " * " << cx("b.y")
             << ";\n";
        work << "      const double zi = " << cx("a.x") << " * " << cx("b.y") << " - " << cx("a.y") << " * " << cx("b.x")
             << ";\n";
        for (const XTerm& T : Gp.terms) {
            std::string c = G(T.gi);
            if (T.zyo) {
                const std::string nm = "u" + std::to_string(tmp++);
                tilec << "    const double " << nm << " = " << tsign(T.zyo) << " ? -" << c << " : " << c << ";\n";
                c = nm;
            }
            if (T.zyr) c = "((__popc(e & " + std::to_string(T.zyr) + "u) & 1) ? -" + c + " : " + c + ")";
            work << "      acc = __builtin_fma(" << (T.im ? "zi" : "zr") << ", " << c << ", acc);\n";
        }
        work << "    }\n";
        body << "    {\n" << tilec.str() << work.str() << "    }\n";
    }
    auto tbase = [&](const std::string& t) {
        std::ostringstream o;
        o << "u64 b_ = (" << t << " << " << L << ");";
        for (int p : lay.hpos) o << " b_ = ((b_ >> " << p << ") << " << (p + 1) << ") | (b_ & " << hex((1ull << p) - 1) << ");";
        return o.str();
    };
    std::ostringstream s;  // hipRTC declares the HIP device API itself: no includes
    s << "typedef unsigned long long u64;\n";
    s << "typedef " << S << " nv2 __attribute__((ext_vector_type(2)));\n";
    s << "__device__ __forceinline__ unsigned swz(unsigned e) { unsigned f = 0;";
    for (int d = digit_bits(dtype); d < K; d += digit_bits(dtype)) s << " f ^= e >> " << d << ";";
    s << " return e ^ (f & " << ((1u << digit_bits(dtype)) - 1u) << "u); }\n";
    s << "__device__ __forceinline__ " << V << " ld(const " << V
      << "* p) { nv2 v = __builtin_nontemporal_load(reinterpret_cast<const nv2*>(p)); " << V
      << " r; r.x = v.x; r.y = v.y; return r; }\n";
    s << "extern \"C\" __global__ void __launch_bounds__(" << NT << ", 1)\n";
    s << "lq_expect(const " << V << "* __restrict__ psi, const double* __restrict__ G, "
      << "double* __restrict__ partial, u64 inst_stride, u64 coef_stride) {\n";
    s << "  __shared__ __align__(16) " << V << " tile[" << (1u << K) << "];\n";
    s << "  const unsigned tid = threadIdx.x;\n";
    s << "  psi += (u64)blockIdx.y * inst_stride;\n  G += (u64)blockIdx.y * coef_stride;\n";
    // each block walks a contiguous range of tiles: consecutive tiles share
    // their pages, which keeps a CU's address translations few (strided
    // ranges measured slower)
    s << "  const u64 per = (" << ntiles << "ull + gridDim.x - 1) / gridDim.x;\n";
    s << "  const u64 t0 = (u64)blockIdx.x * per, tstep = 1;\n";
    s << "  const u64 t1 = t0 + per < " << ntiles << "ull ? t0 + per : " << ntiles << "ull;\n";
    s << "  u64 toff = 0;\n";  // local element e = tid + j * NT -> global offset
    for (int b = 0; b < NTL; ++b) s << "  toff |= (u64)((tid >> " << b << ") & 1u) << " << gpos[b] << ";\n";
    s << "  const unsigned stid = swz(tid);\n";
    s << pre.str();
    s << "  " << V << " r[" << A << "];\n";
    s << "  double acc = 0.0;\n";
    s << "  if (t0 < t1) {\n    " << tbase("t0") << "\n";
    for (int j = 0; j < A; ++j)
        s << "    r[" << j << "] = ld(psi + (b_ | toff | " << hex(spread_local((uint32_t)j << NTL)) << "));\n";
    s << "  }\n";
    s << "  for (u64 t = t0; t < t1; t += tstep) {\n";
    s << "    " << tbase("t") << "\n    const u64 base = b_;\n";
    s << "    __syncthreads();\n";
    for (int j = 0; j < A; ++j) s << "    tile[stid ^ " << swz((uint32_t)j << NTL) << "u] = r[" << j << "];\n";
    s << "    __syncthreads();\n";
    s << "    if (t + tstep < t1) {\n      const u64 tn = t + tstep;\n      " << tbase("tn") << "\n";
    for (int j = 0; j < A; ++j)
        s << "      r[" << j << "] = ld(psi + (b_ | toff | " << hex(spread_local((uint32_t)j << NTL)) << "));\n";
    s << "    }\n";
    s << body.str();
    s << "  }\n";
    s << "  __syncthreads();\n";
    s << "  double* sh = reinterpret_cast<double*>(tile);\n";
    s << "  sh[tid] = acc;\n  __syncthreads();\n";
    s << "  for (unsigned h = " << NT / 2 << "u; h > 0; h >>= 1) { if (tid < h) sh[tid] += sh[tid + h]; __syncthreads(); }\n";
    s << "  if (tid == 0) partial[(u64)blockIdx.y * gridDim.x + blockIdx.x] = sh[0];\n";
    s << "}\n";
    if (derived_out) *derived_out = derived;
    return s.str();
}

bool xk_load(XKernel* k) {
    if (k->fn) return true;
    if (k->failed || k->code.empty()) return false;
    if (hipModuleLoadData(&k->mod, k->code.data()) != hipSuccess ||
        hipModuleGetFunction(&k->fn, k->mod, "lq_expect") != hipSuccess) {
        (void)hipGetLastError();
        k->failed = true;
        k->fn = nullptr;
        jit_count(0, 1, 0.0);
        return false;
    }
    return true;
}

}  // namespace

// Attach kernels to the plan's layouts, compiling the structures the policy
// selects (in parallel); true when every layout has a loaded kernel.
// force: compile at first sight whatever the policy.  load=false: compile
// only (no device).
bool prepare_expect_jit(XPlan& xp, bool force, bool load) {
    const int n = xp.n, dtype = xp.dtype;
    if (xp.layouts.empty()) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (load && (hipStreamIsCapturing(stream(), &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)) force = false;
    const bool always = force || jit_mode() == LQ_JIT_ALWAYS;
    std::vector<std::pair<const std::string*, XKernel*>> todo;
    for (XLayout& lay : xp.layouts) {
        lay.k = nullptr;
        std::string src = layout_source(lay, n, dtype, &lay.derived);
        auto it = g_xk.find(src);
        if (it == g_xk.end()) {
            if (g_xk.size() >= kXMaxKernels) return false;
            it = g_xk.emplace(std::move(src), std::make_unique<XKernel>()).first;
        }
        XKernel* k = it->second.get();
        ++k->seen;
        lay.k = k;
        if (k->failed) return false;
        if (!k->fn && k->code.empty() && (always || k->seen >= 2) && cs == hipStreamCaptureStatusNone &&
            std::find_if(todo.begin(), todo.end(), [&](const auto& e) { return e.second == k; }) == todo.end())
            todo.emplace_back(&it->first, k);
    }
    if (!todo.empty()) {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<char> ok(todo.size(), 0);
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const size_t nth = std::min<size_t>({todo.size(), (size_t)hw, (size_t)16});
        std::vector<std::thread> th;
        for (size_t t = 0; t < nth; ++t)
            th.emplace_back([&, t] {
                for (size_t i = t; i < todo.size(); i += nth) ok[i] = jit_compile_text(*todo[i].first, todo[i].second->code);
            });
        for (auto& x : th) x.join();
        int good = 0, bad = 0;
        for (size_t i = 0; i < todo.size(); ++i) {
            if (ok[i]) {
                ++good;
            } else {
                todo[i].second->failed = true;
                todo[i].second->code.clear();
                ++bad;
            }
        }
        jit_count(good, bad, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    for (XLayout& lay : xp.layouts) {
        if (lay.k->code.empty() && !lay.k->fn) return false;
        if (load && !xk_load(lay.k)) return false;
    }
    return true;
}

// A layout's table: its raw coefficients, then the derived sums.
static void layout_table(const XLayout& l, const std::vector<Combo>& derived, std::vector<double>& out) {
    out.assign(l.coef.begin(), l.coef.end());
    for (const Combo& c : derived) {
        double v = 0.0;
        for (auto& [i, sg] : c) v += sg > 0 ? l.coef[i] : -l.coef[i];
        out.push_back(v);
    }
}

size_t expect_jit_coef_count(const XPlan& xp) {
    size_t c = 0;
    for (const XLayout& l : xp.layouts) c += l.coef.size() + l.derived.size();
    return c;
}

void expect_jit_coefs(const XPlan& xp, double* out) {
    std::vector<double> t;
    for (const XLayout& l : xp.layouts) {
        layout_table(l, l.derived, t);
        std::copy(t.begin(), t.end(), out);
        out += t.size();
    }
}

int expect_jit_blocks(const XPlan& xp, uint64_t ninst) {
    if (!g_cu_count) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&g_cu_count, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            g_cu_count = 256;
    }
    const uint64_t ntiles = 1ull << (xp.n - xp.layouts[0].K);
    const uint64_t want = std::max<uint64_t>(1, (uint64_t)g_cu_count / std::max<uint64_t>(1, ninst));
    return (int)std::min<uint64_t>(ntiles, want);
}

std::string describe_expect_jit(const XPlan& xp) {
    auto esc = [](const std::string& t) {
        std::string o;
        for (char c : t) {
            if (c == '"' || c == '\\') o += '\\', o += c;
            else if (c == '\n') o += "\\n";
            else o += c;
        }
        return o;
    };
    std::ostringstream s;
    s.precision(17);
    const int gx = xp.layouts.empty() ? 0 : expect_jit_blocks(xp, 1);
    s << "{\"n\": " << xp.n << ", \"dtype\": " << xp.dtype << ", \"diag_in\": " << (xp.diag_in ? "true" : "false")
      << ", \"blocks\": " << gx << ", \"rest\": [";
    for (size_t i = 0; i < xp.rest.size(); ++i) s << (i ? ", " : "") << xp.rest[i];
    s << "], \"layouts\": [";
    for (size_t li = 0; li < xp.layouts.size(); ++li) {
        const XLayout& l = xp.layouts[li];
        const uint64_t ntiles = 1ull << (xp.n - l.K);
        s << (li ? ", " : "") << "{\"K\": " << l.K << ", \"L\": " << l.L << ", \"threads\": "
          << (1 << (l.K - window_bits(xp.dtype))) << ", \"ntiles\": " << ntiles << ", \"windows\": " << l.windows.size() << ", \"pair_groups\": " << l.pairs.size()
          << ", \"coef\": [";
        std::vector<Combo> derived;
        const std::string src = layout_source(l, xp.n, xp.dtype, &derived);
        std::vector<double> table;
        layout_table(l, derived, table);
        for (size_t i = 0; i < table.size(); ++i) s << (i ? ", " : "") << table[i];
        s << "], \"src\": \"" << esc(src) << "\"}";
    }
    s << "]}";
    return s.str();
}

// One launch per layout: partials [ninst][gx] per layout, layouts after
// each other in d_partial; coefficients of instance i at d_coef + i *
// coef_stride, layouts after each other.
int launch_expect_jit(const XPlan& xp, const void* psi, const double* d_coef, uint64_t coef_stride,
                      uint64_t ninst, uint64_t inst_stride, int gx, double* d_partial) {
    const double* coef = d_coef;
    double* part = d_partial;
    for (const XLayout& lay : xp.layouts) {
        const void* p = psi;
        void* args[] = {(void*)&p, (void*)&coef, (void*)&part, (void*)&inst_stride, (void*)&coef_stride};
        const int threads = 1 << (lay.K - window_bits(xp.dtype));
        hipError_t e = hipModuleLaunchKernel(lay.k->fn, (unsigned)gx, (unsigned)ninst, 1, (unsigned)threads, 1, 1, 0,
                                             stream(), args, nullptr);
        if (e != hipSuccess) {
            set_error("specialised expectation launch failed: %s", hipGetErrorString(e));
            return LQ_ERR_DEVICE;
        }
        coef += lay.coef.size() + lay.derived.size();
        part += (uint64_t)gx * ninst;
    }
    return LQ_OK;
}

}  // namespace lq
