// Tile-pass kernels specialised at run time.
//
// The generic grouped kernel (k_tileg) reads every op's kind, subcube
// position and control masks at run time, so each op costs the selects and
// path merges that cover all of its cases.  Here the host writes one kernel
// per pass *structure* (tile geometry, group positions, and each op's kind,
// positions and control masks; not the matrices), compiles it with hipRTC
// and launches it in place of k_tileg: every op becomes straight-line
// arithmetic on fixed registers.  Matrices still come from the op table, so
// one compiled kernel serves every parameter value and every instance of a
// batch.
//
// Policy (lq_set_jit): off; "repeat" (default) compiles a structure the
// second time a plan holds it, and only for passes of at least
// kJitMinTiles tiles (smaller passes are launch-bound); "always" compiles
// everything at first sight.  Compiles of one plan run in parallel threads.
// A failed compile leaves that structure on the generic kernel for good.
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lq_fused.h"
#include "lq_host.h"

// kTileTypesSrc / kTileOpsSrc: the text of lq_tile_types.h / lq_tile_ops.h,
// generated by the build next to the object files.
#include "lq_jit_src.inc"

namespace lq {

struct JitKernel {
    int seen = 0;
    bool failed = false;
    std::vector<char> code;  // code object from hipRTC
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
};

static int g_jit_mode = LQ_JIT_REPEAT;
static constexpr uint64_t kJitMinTiles = 1024;
static constexpr uint64_t kStreamBytes = 1ull << 30;
// First / last groups of a pass straight from / to the state (on by default;
// LQ_JIT_DIRECT=0 in the environment keeps every pass on LDS tile loads and
// stores, for comparison)
static const bool direct_groups = !(std::getenv("LQ_JIT_DIRECT") && std::getenv("LQ_JIT_DIRECT")[0] == '0');
// Z / CZ as sign flips (LQ_JIT_SIGN=0: as complex phases like other diagonals)
static const bool sign_flips = !(std::getenv("LQ_JIT_SIGN") && std::getenv("LQ_JIT_SIGN")[0] == '0');
static constexpr size_t kJitMaxStructures = 4096;
static std::unordered_map<std::string, std::unique_ptr<JitKernel>> g_jit;
static long long g_jit_compiled = 0, g_jit_failed = 0, g_jit_launches = 0;
static double g_jit_seconds = 0.0;

static constexpr int jit_threads(int K) { return K <= 12 ? 256 : (1 << (K - 4)); }

// Kernel text for one pass: constants K, HB, the groups' subcube positions
// and every op's kind, positions and masks.  Op i of the pass reads its
// matrix from ops[i].
//
// Within a group, three rewrites (all exact up to rounding):
// * X gates with their controls outside the subcube and nothing but other
//   such gates on their bit before (after) them in the group are done by
//   flipping the gather (scatter) slots of the threads whose controls are
//   set, instead of a per-lane register permutation.
// * Diagonal ops that depend on subcube bits only are moved past the
//   non-diagonal ops that do not act on those bits, and each run of two or
//   more becomes one 16-entry phase table per tile: sixteen threads form the
//   products (in f64) into LDS while the tile loads, and the group applies
//   the whole run with one complex multiply per amplitude.
// * Sign ops (Z, CZ: kTDiagNeg) move the same way and are applied as sign
//   flips: an op whose controls lie inside the subcube negates its slots at
//   compile time; one with controls outside flips the sign bits of its slots
//   by a per-thread mask (an XOR, no multiply and no table entry).
static constexpr int kMaxRuns = 32;

struct GroupPlan {
    std::vector<int> fold;            // per op of the group: 0, 1 gather, 2 scatter
    // (0, op index), (1, run id), (2, 0: the thread phase) or (3, sign set id), in order
    std::vector<std::pair<int, int>> items;
    std::vector<int> uphase;          // diagonal ops on outside bits only, merged into one thread phase
    std::vector<int> uneg;            // sign ops on outside bits only, folded into the thread phase
    std::vector<std::vector<int>> signs;  // sign sets: ops applied together as sign flips
};

// init: the pass starts from |0...0> (j_tile_zero in place of the tile
// load), for a circuit's first pass over a state not yet written.  stream:
// nontemporal global loads and stores (states far beyond the caches).
static std::string pass_source(const TilePass& p, const TileOp* ops, const TileGroup* groups, int dtype,
                               bool init = false, bool stream = false) {
    const int K = p.L + p.hb;
    const int NT = jit_threads(K);
    auto touch = [&](const TileOp& t) -> int {
        int m = t.cin;
        if (t.kind == kTDiag) {
            if (t.s0p >= 0) m |= 1 << t.s0p;
            if (t.s1p >= 0) m |= 1 << t.s1p;
        } else {
            m |= 1 << t.sp0;
            if (t.kind == kTSwap2 || t.kind == kTQuad) m |= 1 << t.sp1;
        }
        return m;
    };
    auto foldable = [&](const TileOp& t) { return t.kind == kTPairSwap && t.cin == 0 && t.cout != 0; };
    // diagonal and depending on subcube bits only
    auto local_diag = [&](const TileOp& t) {
        return t.kind == kTDiag && t.cout == 0 && t.osel0 < 0 && t.osel1 < 0;
    };
    // diagonal and depending on bits outside the subcube only
    auto outside_diag = [&](const TileOp& t) {
        return t.kind == kTDiag && t.cin == 0 && t.s0p < 0 && t.s1p < 0 && (t.cout != 0 || t.osel0 >= 0 || t.osel1 >= 0);
    };
    // negates where its controls are set (checked on the values as well)
    auto sign_op = [&](const TileOp& t) {
        return sign_flips && t.kind == kTDiag && t.dclass == kTDiagNeg && t.sel0 < 0 && t.sel1 < 0 && t.ctrl != 0 &&
               t.m[0] == -1.0 && t.m[1] == 0.0;
    };
    std::vector<std::vector<int>> runs;  // op indices (pass-relative) of each merged run
    std::vector<GroupPlan> gp(p.group_count);
    for (int gi = 0; gi < p.group_count; ++gi) {
        const TileGroup& G = groups[p.group_first + gi];
        const int o0 = G.op_first, o1 = G.op_first + G.op_count;
        GroupPlan& P = gp[gi];
        P.fold.assign(G.op_count, 0);
        int seen = 0;
        for (int o = o0; o < o1; ++o) {
            const TileOp& t = ops[p.op_first + o];
            if (foldable(t) && !(seen & (1 << t.sp0))) P.fold[o - o0] = 1;
            else seen |= touch(t);
        }
        seen = 0;
        for (int o = o1 - 1; o >= o0; --o) {
            const TileOp& t = ops[p.op_first + o];
            if (P.fold[o - o0]) continue;
            if (foldable(t) && !(seen & (1 << t.sp0))) P.fold[o - o0] = 2;
            else seen |= touch(t);
        }
        // pending local diagonals and sign ops sink past non-diagonal ops on
        // other bits
        std::vector<int> pend, pend_sign;
        int pend_dep = 0;
        auto flush = [&] {
            if (pend.size() >= 2 && (int)runs.size() < kMaxRuns) {
                P.items.emplace_back(1, (int)runs.size());
                runs.push_back(pend);
            } else {
                for (int o : pend) P.items.emplace_back(0, o);
            }
            if (!pend_sign.empty()) {
                P.items.emplace_back(3, (int)P.signs.size());
                P.signs.push_back(pend_sign);
            }
            pend.clear();
            pend_sign.clear();
            pend_dep = 0;
        };
        for (int o = o0; o < o1; ++o) {
            if (P.fold[o - o0]) continue;
            const TileOp& t = ops[p.op_first + o];
            if (sign_op(t) && t.cin == 0) {
                P.uneg.push_back(o);  // commutes with every op of the group
            } else if (sign_op(t)) {
                pend_sign.push_back(o);
                pend_dep |= touch(t);
            } else if (outside_diag(t)) {
                P.uphase.push_back(o);  // commutes with every op of the group
            } else if (local_diag(t)) {
                pend.push_back(o);
                pend_dep |= touch(t);
            } else if (t.kind == kTDiag) {
                P.items.emplace_back(0, o);  // diagonals commute with each other
            } else {
                int sup = 1 << t.sp0;
                if (t.kind == kTSwap2 || t.kind == kTQuad) sup |= 1 << t.sp1;
                if (sup & pend_dep) flush();
                P.items.emplace_back(0, o);
            }
        }
        flush();
        // outside sign ops: into the thread phase when there is one, else
        // into a sign set (their flag flips all sixteen slots)
        if (P.uphase.size() + (P.uphase.empty() ? 0 : P.uneg.size()) >= 2) {
            P.items.emplace_back(2, 0);
        } else {
            for (int o : P.uphase) P.items.emplace_back(0, o);
            P.uphase.clear();
            if (!P.uneg.empty()) {
                if (P.signs.empty()) {
                    P.items.emplace_back(3, 0);
                    P.signs.emplace_back();
                }
                P.signs.back().insert(P.signs.back().end(), P.uneg.begin(), P.uneg.end());
                P.uneg.clear();
            }
        }
    }

    // The first (last) group reads (writes) the state directly instead of
    // through a tile load (store) in LDS when its subcube bits leave the
    // lowest bits of the tile to the lanes, so that a wave's access still
    // covers whole 128-byte lines; a pass of one such group touches no LDS.
    const int lmin = dtype == LQ_C128 ? 3 : 4;
    auto lanes_low = [&](int gi) {
        const TileGroup& G = groups[p.group_first + gi];
        int m = K;
        for (int k = 0; k < kGroupBits; ++k) m = std::min(m, (int)G.spos[k]);
        return p.L >= lmin && m >= lmin;
    };
    const bool dfirst = direct_groups && p.group_count >= 1 && lanes_low(0);
    const bool single = dfirst && p.group_count == 1;
    const bool dlast = direct_groups && p.group_count >= 2 && lanes_low(p.group_count - 1);
    // a direct first group that reads the state is issued one tile ahead
    // when the registers allow: at two waves per SIMD (256 VGPRs each:
    // blocks of 256 threads, or complex128 2^13 tiles whose 128 KiB of LDS
    // leave one block of 512 threads per CU; brickwork 264 -> 240 ms at
    // 13 / 5 bits); complex64 2^13 tiles run two such blocks per CU, four
    // waves per SIMD, where the extra 16 amplitudes per thread cost more
    // than the overlap gains (complex64 brickwork 127 -> 138 ms)
    const bool pipe = dfirst && !init && (NT <= 256 || dtype == LQ_C128);
    // a pass of one direct group issues the next tile's gather before its
    // stores (the vector memory counter retires in order, so waiting for
    // loads issued after stores waits for the stores too); in longer passes
    // the gather follows the first group (issuing it before the group costs
    // 64 register moves per tile and gained nothing)
    const bool early = single;
    // gather (which = 1) / scatter (2) slot flips of group gi, as a lambda's text
    auto flips_of = [&](int gi, int which) {
        const TileGroup& G = groups[p.group_first + gi];
        const GroupPlan& P = gp[gi];
        const int o0 = G.op_first, o1 = G.op_first + G.op_count;
        std::ostringstream f;
        f << "[&](uint64_t gb) -> uint32_t { uint32_t f = 0u;";
        for (int o = o0; o < o1; ++o) {
            if (P.fold[o - o0] != which) continue;
            const TileOp& t = ops[p.op_first + o];
            f << " if ((gb & 0x" << std::hex << t.cout << "ull) == 0x" << t.cout << std::dec << "ull) f ^= "
              << (1u << G.spos[t.sp0]) << "u;";
        }
        f << " return f; }";
        return f.str();
    };
    auto pos_of = [&](int gi) {
        const TileGroup& G = groups[p.group_first + gi];
        return std::to_string(G.spos[0]) + ", " + std::to_string(G.spos[1]) + ", " + std::to_string(G.spos[2]) +
               ", " + std::to_string(G.spos[3]);
    };
    std::ostringstream s;
    s << "#include \"lq_tile_ops.h\"\nusing namespace lq;\n";
    if (dtype == LQ_C128)
        s << "typedef double T;\ntypedef double V __attribute__((ext_vector_type(2)));\n";
    else
        s << "typedef float T;\ntypedef float V __attribute__((ext_vector_type(2)));\n";
    s << "constexpr int K = " << K << ", HB = " << p.hb << ", NT = " << NT << ";\n";
    s << "extern \"C\" __global__ void __launch_bounds__(NT, 2) lq_pass(V* __restrict__ psi, TilePass p, "
         "const TileOp* __restrict__ ops, uint64_t ninst, uint64_t inst_stride, uint64_t ops_stride) {\n";
    s << "  __shared__ __attribute__((aligned(16))) V tile[1u << K];\n";
    s << "  __shared__ uint64_t hoff[1u << HB];\n";
    if (!runs.empty()) s << "  __shared__ __attribute__((aligned(16))) V ptab[" << runs.size() * 16 << "];\n";
    s << "  j_pass_begin<K, HB, NT>(hoff, p);\n";
    // each block walks a contiguous range of tiles: consecutive tiles share
    // an instance, so its op table stays in the scalar cache
    // p.ntiles is a power of two: instance and tile index by shift and mask
    s << "  const uint64_t total = ninst * p.ntiles, tmask = p.ntiles - 1ull;\n";
    s << "  const uint32_t tsh = (uint32_t)__builtin_ctzll(p.ntiles);\n";
    s << "  const uint64_t per = (total + gridDim.x - 1) / gridDim.x;\n";
    s << "  const uint64_t w0 = (uint64_t)blockIdx.x * per, w1 = w0 + per < total ? w0 + per : total;\n";
    if (!runs.empty()) s << "  uint32_t tab_inst = 0xffffffffu;\n";
    const char* sm = stream ? "true" : "false";
    if (pipe) {
        s << "  V pre[16];\n  const auto gf0 = " << flips_of(0, 1) << ";\n";
        s << "  if (w0 < w1) {\n    const uint32_t in0 = __builtin_amdgcn_readfirstlane((uint32_t)(w0 >> tsh));\n";
        s << "    j_first_fetch<K, HB, NT, " << pos_of(0) << ", " << sm
          << ">(pre, psi + (uint64_t)in0 * inst_stride, j_tile_base<K, HB>(w0 & tmask, p), hoff, p, "
             "gf0);\n  }\n";
    }
    s << "  for (uint64_t w = w0; w < w1; ++w) {\n";
    s << "    const uint32_t inst = __builtin_amdgcn_readfirstlane((uint32_t)(w >> tsh));\n";
    s << "    const uint64_t base = j_tile_base<K, HB>(w & tmask, p);\n";
    s << "    V* st = psi + (uint64_t)inst * inst_stride;\n";
    s << "    const TileOp* op = ops + (uint64_t)inst * ops_stride + p.op_first;\n";
    if (!runs.empty()) {
        // phase tables: run r by the sixteen threads of slice r mod NT/16,
        // when the block reaches a new instance (its tiles are contiguous,
        // so mostly once); the tile load's barrier publishes them, and the
        // previous tile's last barrier has retired every read of the old ones
        s << "    if (inst != tab_inst) {\n      tab_inst = inst;\n";
        // with a first group read straight from the state, no tile load's
        // barrier follows: barriers of their own retire the previous tile's
        // reads of the old tables and publish the new ones
        if (dfirst) s << "      __syncthreads();\n";
        s << "      const int ts = (int)(threadIdx.x & 15u), tg = (int)(threadIdx.x >> 4);\n";
        for (size_t r = 0; r < runs.size(); ++r) {
            s << "      if (tg == " << (int)(r % (size_t)(NT / 16)) << ") {\n        double pr = 1.0, pi = 0.0;\n";
            for (int o : runs[r]) {
                const TileOp& t = ops[p.op_first + o];
                s << "        j_phase_mul<" << t.s0p << ", " << t.s1p << ", " << t.cin << ">(pr, pi, ts, op[" << o
                  << "]);\n";
            }
            s << "        ptab[" << r * 16 << " + ts] = V{(T)pr, (T)pi};\n      }\n";
        }
        if (dfirst) s << "      __syncthreads();\n";
        s << "    }\n";
    }
    if (!dfirst) {
        if (init) s << "    j_tile_zero<K, NT>(tile, base);\n";
        else s << "    j_tile_load<K, HB, NT, " << sm << ">(tile, st, base, hoff);\n";
    }
    for (int gi = 0; gi < p.group_count; ++gi) {
        const TileGroup& G = groups[p.group_first + gi];
        const GroupPlan& P = gp[gi];
        auto flips = [&](int which) { return flips_of(gi, which); };
        const std::string pos = std::to_string(G.spos[0]) + ", " + std::to_string(G.spos[1]) + ", " +
                                std::to_string(G.spos[2]) + ", " + std::to_string(G.spos[3]);
        if (gi == 0 && pipe)
            s << "    j_first_apply<K, HB, NT, " << pos << ", " << sm << ", " << (single ? "true" : "false") << ", "
              << (early ? "true" : "false") << ", T>(tile, st, pre, base, hoff, p, [&](T* xr, T* xi, uint64_t gb) {\n";
        else if (gi == 0 && dfirst)
            s << "    j_group_first<K, HB, NT, " << pos << ", " << sm << ", " << (init ? "true" : "false") << ", "
              << (single ? "true" : "false") << ", T>(tile, st, base, hoff, p, [&](T* xr, T* xi, uint64_t gb) {\n";
        else if (gi == p.group_count - 1 && dlast)
            s << "    j_group_last<K, HB, NT, " << pos << ", " << sm << ", " << (dfirst ? "false" : "true")
              << ", T>(tile, st, base, hoff, p, [&](T* xr, T* xi, uint64_t gb) {\n";
        else
            s << "    j_group<K, HB, NT, " << pos << ", T>(tile, base, hoff, [&](T* xr, T* xi, uint64_t gb) {\n";
        auto gbit = [](int b) { return "(int)((gb >> " + std::to_string(b) + ") & 1ull)"; };
        for (const auto& [what, idx] : P.items) {
            s << "      ";
            if (what == 2) {
                s << "{\n        double pr = 1.0, pi = 0.0;\n";
                for (int o : P.uphase) {
                    const TileOp& t = ops[p.op_first + o];
                    std::ostringstream on, k;
                    on << "(gb & 0x" << std::hex << t.cout << "ull) == 0x" << t.cout << std::dec << "ull";
                    k << (t.osel0 >= 0 ? gbit(t.osel0) : std::string("0")) << " | ("
                      << (t.osel1 >= 0 ? gbit(t.osel1) : std::string("0")) << " << 1)";
                    s << "        j_uphase_mul(pr, pi, " << on.str() << ", " << k.str() << ", op[" << o << "]);\n";
                }
                for (int o : P.uneg) {
                    const TileOp& t = ops[p.op_first + o];
                    s << "        j_uphase_neg(pr, pi, (gb & 0x" << std::hex << t.cout << "ull) == 0x" << t.cout
                      << std::dec << "ull);\n";
                }
                s << "        j_uphase_apply<T>(xr, xi, pr, pi);\n      }\n";
                continue;
            }
            if (what == 3) {
                // slot c is negated by every op of the set whose in-subcube
                // controls it has: at compile time for ops with no outside
                // controls, else by the op's per-thread sign-bit flag
                // (the flags of ops on outside bits only, which flip every
                // slot, are combined first into nu)
                const std::vector<int>& set = P.signs[idx];
                s << "{\n";
                std::vector<int> flag(set.size(), -1);
                int nflags = 0;
                std::string nu;
                for (size_t k = 0; k < set.size(); ++k) {
                    const TileOp& t = ops[p.op_first + set[k]];
                    if (!t.cout) continue;
                    flag[k] = nflags++;
                    s << "        const uint32_t n" << flag[k] << " = (gb & 0x" << std::hex << t.cout << "ull) == 0x"
                      << t.cout << std::dec << "ull ? 0x80000000u : 0u;\n";
                    if (!t.cin) nu += (nu.empty() ? "n" : " ^ n") + std::to_string(flag[k]);
                }
                if (!nu.empty()) s << "        const uint32_t nu = " << nu << ";\n";
                for (int c = 0; c < 16; ++c) {
                    bool neg = false;
                    std::string f = nu.empty() ? "" : "nu";
                    for (size_t k = 0; k < set.size(); ++k) {
                        const TileOp& t = ops[p.op_first + set[k]];
                        if ((c & t.cin) != t.cin || (flag[k] >= 0 && !t.cin)) continue;
                        if (flag[k] < 0) neg = !neg;
                        else f += (f.empty() ? "n" : " ^ n") + std::to_string(flag[k]);
                    }
                    if (neg) s << "        xr[" << c << "] = -xr[" << c << "]; xi[" << c << "] = -xi[" << c << "];\n";
                    if (!f.empty()) s << "        j_flip(xr[" << c << "], xi[" << c << "], " << f << ");\n";
                }
                s << "      }\n";
                continue;
            }
            if (what == 1) {
                int dep = 0;
                for (int o : runs[idx]) dep |= touch(ops[p.op_first + o]);
                s << "j_diag_tab<" << dep << ", T>(xr, xi, ptab + " << idx * 16 << ");\n";
                continue;
            }
            const int o = idx;
            const TileOp& t = ops[p.op_first + o];
            if (t.cout) s << "if ((gb & 0x" << std::hex << t.cout << "ull) == 0x" << t.cout << std::dec << "ull) ";
            switch (t.kind) {
                case kTDiag: {
                    auto tb = [&](int osel) {
                        std::string r = "0";
                        if (osel >= 0) r = "(int)((gb >> " + std::to_string(osel) + ") & 1ull)";
                        return r;
                    };
                    s << "j_diag<" << t.s0p << ", " << t.s1p << ", " << t.cin << ", T>(xr, xi, op[" << o << "], "
                      << tb(t.osel0) << ", " << tb(t.osel1) << ");\n";
                    break;
                }
                case kTSwap2:
                    s << "g_swap2<" << t.sp0 << ", " << t.sp1 << ", T>(xr, xi);\n";
                    break;
                case kTQuad:
                    s << "g_quad<" << t.sp0 << ", " << t.sp1 << ", T>(xr, xi, op[" << o << "].m);\n";
                    break;
                default:
                    s << "j_pair<" << t.sp0 << ", " << t.kind << ", " << t.cin << ", T>(xr, xi, op[" << o << "]);\n";
                    break;
            }
        }
        if (gi == 0 && pipe) {
            // the next tile's first gather, issued once this tile's is in
            // registers: in flight while this tile's groups run
            s << "    }, " << flips(2) << ", [&] {\n";
            s << "      if (w + 1 < w1) {\n        const uint32_t in1 = __builtin_amdgcn_readfirstlane((uint32_t)((w + 1) >> tsh));\n";
            s << "        j_first_fetch<K, HB, NT, " << pos << ", " << sm
              << ">(pre, psi + (uint64_t)in1 * inst_stride, j_tile_base<K, HB>((w + 1) & tmask, p), "
                 "hoff, p, gf0);\n      }\n    });\n";
        } else {
            s << "    }, " << flips(1) << ", " << flips(2) << ");\n";
        }
    }
    if (!dlast && !single) s << "    j_tile_store<K, HB, NT, " << sm << ">(tile, st, base, hoff);\n";
    s << "  }\n}\n";
    return s.str();
}

static const char* kJitOpts[] = {"--offload-arch=gfx950", "-O3", "-ffp-contract=fast",
                                 "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                                 "-mllvm", "-simplifycfg-sink-common=false"};

// Compile one kernel text; false (with the log on stderr when LQ_JIT_LOG is
// set) on failure.
static bool jit_compile(const std::string& src, std::vector<char>& code) {
    if (const char* dir = std::getenv("LQ_JIT_DUMP")) {
        // the kernel text, for reading (LQ_JIT_DUMP=<directory>)
        const std::string path = std::string(dir) + "/lq_pass_" + std::to_string(std::hash<std::string>{}(src)) + ".hip";
        if (FILE* f = std::fopen(path.c_str(), "w")) {
            std::fputs(src.c_str(), f);
            std::fclose(f);
        }
    }
    const char* headers[] = {kTileTypesSrc, kTileOpsSrc};
    const char* names[] = {"lq_tile_types.h", "lq_tile_ops.h"};
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, src.c_str(), "lq_pass.hip", 2, headers, names) != HIPRTC_SUCCESS) return false;
    hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(kJitOpts) / sizeof(kJitOpts[0])), kJitOpts);
    bool ok = r == HIPRTC_SUCCESS;
    if (!ok && std::getenv("LQ_JIT_LOG")) {
        size_t n = 0;
        if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
            std::string log(n, '\0');
            if (hiprtcGetProgramLog(prog, &log[0]) == HIPRTC_SUCCESS) std::fprintf(stderr, "%s\n", log.c_str());
        }
    }
    size_t n = 0;
    if (ok) ok = hiprtcGetCodeSize(prog, &n) == HIPRTC_SUCCESS && n > 0;
    if (ok) {
        code.resize(n);
        ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
    }
    (void)hiprtcDestroyProgram(&prog);
    return ok;
}

int jit_mode() { return g_jit_mode; }

bool jit_compile_text(const std::string& src, std::vector<char>& code) { return jit_compile(src, code); }

void jit_count(int compiled, int failed, double seconds) {
    g_jit_compiled += compiled;
    g_jit_failed += failed;
    g_jit_seconds += seconds;
}

// Compile the given entries in parallel (at most 16 threads).
static void compile_all(const std::vector<std::pair<const std::string*, JitKernel*>>& todo) {
    if (todo.empty()) return;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<char> ok(todo.size(), 0);
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nth = std::min<size_t>({todo.size(), (size_t)hw, (size_t)16});
    std::vector<std::thread> th;
    for (size_t t = 0; t < nth; ++t)
        th.emplace_back([&, t] {
            for (size_t i = t; i < todo.size(); i += nth) ok[i] = jit_compile(*todo[i].first, todo[i].second->code);
        });
    for (auto& x : th) x.join();
    for (size_t i = 0; i < todo.size(); ++i) {
        if (ok[i]) {
            ++g_jit_compiled;
        } else {
            todo[i].second->failed = true;
            todo[i].second->code.clear();
            ++g_jit_failed;
        }
    }
    g_jit_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Load a compiled code object on the current device.
static bool jit_load(JitKernel* k) {
    if (k->fn) return true;
    if (k->failed || k->code.empty()) return false;
    if (hipModuleLoadData(&k->mod, k->code.data()) != hipSuccess ||
        hipModuleGetFunction(&k->fn, k->mod, "lq_pass") != hipSuccess) {
        (void)hipGetLastError();
        k->failed = true;
        k->fn = nullptr;
        ++g_jit_failed;
        return false;
    }
    return true;
}

static bool jit_eligible(const TilePass& p, uint64_t ninst, bool force) {
    const int K = p.L + p.hb;
    if (K < 8 || K > 13 || p.group_count <= 0) return false;
    return force || p.ntiles * ninst >= kJitMinTiles;
}

// Attach specialised kernels to a plan's passes (Step::jit), compiling the
// structures the policy selects.  load=false: compile only (no device).
// compile_now: compile every eligible structure at once, whatever the mode
// (explicit requests); any_size: ignore the minimum pass size; init_first:
// the first pass builds |0...0> in LDS instead of reading the state.
static int jit_prepare_impl(Plan& plan, int dtype, uint64_t ninst, bool load, bool compile_now, bool any_size,
                            bool init_first = false) {
    for (Step& s : plan.steps) s.jit = nullptr;
    if (g_jit_mode == LQ_JIT_OFF && !compile_now) return 0;
    const bool always = compile_now || g_jit_mode == LQ_JIT_ALWAYS;
    // the pass's states against the caches (256 MB Infinity Cache): passes
    // over more than kStreamBytes stream through nontemporally
    const double bytes = (double)ninst * std::ldexp(1.0, plan.n) * (dtype == LQ_C128 ? 16.0 : 8.0);
    const bool stream = bytes > (double)kStreamBytes;
    std::vector<std::pair<Step*, JitKernel*>> use;
    std::vector<std::pair<const std::string*, JitKernel*>> todo;
    for (Step& s : plan.steps) {
        if (s.normalize || !jit_eligible(s.pass, ninst, any_size || g_jit_mode == LQ_JIT_ALWAYS)) continue;
        std::string src = pass_source(s.pass, plan.ops.data(), plan.groups.data(), dtype,
                                      init_first && &s == &plan.steps[0], stream);
        if (const char* dir = std::getenv("LQ_JIT_DUMP")) {
            // which kernel text each step of the plan uses, for reading
            if (FILE* f = std::fopen((std::string(dir) + "/steps.txt").c_str(), "a")) {
                std::fprintf(f, "n=%d step %d L=%d hb=%d groups=%d ops=%d lq_pass_%zu hpos", plan.n,
                             (int)(&s - plan.steps.data()), s.pass.L, s.pass.hb, s.pass.group_count,
                             s.pass.op_count, std::hash<std::string>{}(src));
                for (int j = 0; j < s.pass.hb; ++j) std::fprintf(f, " %d", s.pass.hpos[j]);
                std::fputc('\n', f);
                std::fclose(f);
            }
        }
        auto it = g_jit.find(src);
        if (it == g_jit.end()) {
            if (g_jit.size() >= kJitMaxStructures) continue;  // bounded cache: new structures stay generic
            it = g_jit.emplace(std::move(src), std::make_unique<JitKernel>()).first;
        }
        JitKernel* k = it->second.get();
        ++k->seen;
        if (k->failed) continue;
        if (!k->fn && k->code.empty() && (always || k->seen >= 2) &&
            std::find_if(todo.begin(), todo.end(), [&](const auto& e) { return e.second == k; }) == todo.end())
            todo.emplace_back(&it->first, k);
        use.emplace_back(&s, k);
    }
    compile_all(todo);
    int ready = 0;
    for (auto& [s, k] : use) {
        if (k->code.empty() && !k->fn) continue;
        if (load && !jit_load(k)) continue;
        if (load) s->jit = k;
        ++ready;
    }
    return ready;
}

void jit_prepare(Plan& plan, int dtype, uint64_t ninst) {
    // never inside a stream capture: compiling and loading are not stream work
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream(), &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
        for (Step& s : plan.steps) s.jit = nullptr;
        return;
    }
    (void)jit_prepare_impl(plan, dtype, ninst, true, false, false);
}

bool jit_prepare_batch(Plan& plan, int dtype, uint64_t ninst) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream(), &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
        for (Step& s : plan.steps) s.jit = nullptr;
        return false;
    }
    const int ready = jit_prepare_impl(plan, dtype, ninst, true, false, true, true);
    return ready == (int)plan.steps.size();
}

int jit_warm(int n, int dtype, const lq_gate* gates, int count) {
    if (g_jit_mode == LQ_JIT_OFF) return LQ_OK;
    std::vector<Prim> prims;
    int rc = gates_to_prims(n, gates, count, prims);
    if (rc) return rc;
    Plan plan;
    plan_passes(n, prims, default_tile_bits(n, dtype), default_low_bits(n, dtype), plan);
    (void)jit_prepare_impl(plan, dtype, 1, true, true, false);
    return LQ_OK;
}

int launch_jit(const JitKernel* k, void* psi, const TilePass& p, const TileOp* d_ops, uint64_t ninst,
               uint64_t inst_stride, uint64_t ops_stride, unsigned grid) {
    const int K = p.L + p.hb;
    if (p.ntiles == 0 || (p.ntiles & (p.ntiles - 1)) != 0) {
        // the kernels split a work index into instance and tile by shift
        set_error("specialised tile pass: %llu tiles per state is not a power of two",
                  (unsigned long long)p.ntiles);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    TilePass pp = p;
    void* args[] = {&psi, &pp, &d_ops, &ninst, &inst_stride, &ops_stride};
    hipError_t e = hipModuleLaunchKernel(k->fn, grid, 1, 1, (unsigned)jit_threads(K), 1, 1, 0, stream(), args,
                                         nullptr);
    if (e != hipSuccess) {
        set_error("specialised tile pass launch failed: %s", hipGetErrorString(e));
        return LQ_ERR_DEVICE;
    }
    ++g_jit_launches;
    return LQ_OK;
}

}  // namespace lq

using namespace lq;

extern "C" int lq_set_jit(int mode) {
    LQ_API_GUARD;
    clear_error();
    if (mode != LQ_JIT_OFF && mode != LQ_JIT_REPEAT && mode != LQ_JIT_ALWAYS) {
        set_error("jit mode must be LQ_JIT_OFF, LQ_JIT_REPEAT or LQ_JIT_ALWAYS");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    g_jit_mode = mode;
    return LQ_OK;
}

extern "C" int lq_jit_stats(long long* compiled, long long* failed, long long* launches, double* compile_seconds) {
    LQ_API_GUARD;
    clear_error();
    if (compiled) *compiled = g_jit_compiled;
    if (failed) *failed = g_jit_failed;
    if (launches) *launches = g_jit_launches;
    if (compile_seconds) *compile_seconds = g_jit_seconds;
    return LQ_OK;
}

extern "C" int lq_jit_compile_plan(int num_qubits, int dtype, const lq_gate* gates, int count, int* kernels) {
    LQ_API_GUARD;
    clear_error();
    if (num_qubits < 1 || num_qubits > kMaxQubits || count < 0 || (count > 0 && !gates) ||
        (dtype != LQ_C128 && dtype != LQ_C64)) {
        set_error("invalid arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    std::vector<Prim> prims;
    int rc = gates_to_prims(num_qubits, gates, count, prims);
    if (rc) return rc;
    Plan plan;
    plan_passes(num_qubits, prims, default_tile_bits(num_qubits, dtype), default_low_bits(num_qubits, dtype), plan);
    const long long before = g_jit_failed;
    const int ready = jit_prepare_impl(plan, dtype, 1, false, true, true);
    if (kernels) *kernels = ready;
    if (g_jit_failed != before) {
        set_error("hipRTC failed on %lld pass kernel(s) (LQ_JIT_LOG=1 prints the log)", g_jit_failed - before);
        return LQ_ERR_DEVICE;
    }
    return LQ_OK;
}
