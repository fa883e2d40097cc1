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
#include <type_traits>

#include "lq_expect_jit.h"
#include "lq_host.h"
#include "lq_tile_ops.h"

namespace lq {

constexpr int kExpBlocksMax = 8192;
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

// ---- flip groups from LDS tiles ------------------------------------------
// A tile of 2^K amplitudes (the L lowest bits and hb chosen high bits, as in
// the fused engine's passes) is read into LDS once, and every flip group
// whose mask lies inside the tile's bits is evaluated from there: one read
// of the state serves all of them.  For a pair (i, j = i ^ x) and a term t
// with sign s_t(i) = (-1)^{popc(i & zy_t)}, the two rows' contributions add
// up to s_t(i) * g_t * Re(conj(psi_i) psi_j) when popc(x & zy_t) is even
// (g_t = 2 cr_t) and s_t(i) * g_t * Im(conj(psi_i) psi_j) when it is odd
// (g_t = -2 ci_t).  popc(i & zy_t) splits into the tile base's bits outside
// the tile (zyo, uniform per tile) and the local index's bits (zyl).
// 32 KiB tiles (2^11 complex128, 2^12 complex64): four blocks per CU hide
// the LDS latency better than two blocks of 64 KiB tiles, and the smaller
// tiles still cover the usual local Hamiltonians in a few layouts.
template <int K, typename V>
__global__ void __launch_bounds__(kExpThreads, ((sizeof(V) << K) <= 32768 ? 4 : 2))
    k_expect_tile(const V* __restrict__ psi, ExpGeom geo, uint64_t inst_stride,
                  const ExpTileGroup* __restrict__ groups, int ngroups, const ExpTileTerm* __restrict__ terms,
                  double* __restrict__ partial) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr uint32_t TS = 1u << K;
    constexpr int EPT = TS >= (uint32_t)kExpThreads ? (int)(TS / kExpThreads) : 1;
    V* tile = reinterpret_cast<V*>(smem);
    uint64_t* hoff = reinterpret_cast<uint64_t*>(smem + TS * sizeof(V));
    double* sh = reinterpret_cast<double*>(smem);  // reused for the final sum
    const uint32_t tid = threadIdx.x;
    const uint32_t lowmask = (1u << geo.L) - 1u;
    for (uint32_t h = tid; h < (1u << geo.hb); h += kExpThreads) {
        uint64_t o = 0;
        for (int j = 0; j < geo.hb; ++j)
            if ((h >> j) & 1u) o |= 1ull << geo.hpos[j];
        hoff[h] = o;
    }
    __syncthreads();
    const V* s = psi + (uint64_t)blockIdx.y * inst_stride;
    const uint64_t per = (geo.ntiles + gridDim.x - 1) / gridDim.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * per, t1 = t0 + per < geo.ntiles ? t0 + per : geo.ntiles;
    double acc = 0.0;
    for (uint64_t t = t0; t < t1; ++t) {
        uint64_t base = t << geo.L;
        for (int j = 0; j < geo.hb; ++j) base = insert_zero(base, geo.hpos[j]);
        {
            V r[EPT];
            constexpr bool FULLT = TS >= (uint32_t)kExpThreads;
#pragma unroll
            for (int j = 0; j < EPT; ++j) {
                const uint32_t e = tid + (uint32_t)j * kExpThreads;
                if (FULLT || e < TS) r[j] = s[base | (e & lowmask) | hoff[e >> geo.L]];
            }
#pragma unroll
            for (int j = 0; j < EPT; ++j) {
                const uint32_t e = tid + (uint32_t)j * kExpThreads;
                if (FULLT || e < TS) tile[swz(e)] = r[j];
            }
        }
        __syncthreads();
        constexpr int PPT = TS / 2 >= (uint32_t)kExpThreads ? (int)(TS / 2 / kExpThreads) : 1;
        for (int g = 0; g < ngroups; ++g) {
            const ExpTileGroup G = groups[g];
            const uint32_t sxl = swz(G.xl);
            // terms four at a time, their tile-uniform sign folded into the
            // weights (gr for the real part, gi for the imaginary part)
            for (int q0 = G.first; q0 < G.first + G.count; q0 += 4) {
                const int nq = G.first + G.count - q0 < 4 ? G.first + G.count - q0 : 4;
                uint32_t zl[4];
                double gr[4], gi[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    zl[j] = 0u;
                    gr[j] = gi[j] = 0.0;
                    if (j < nq) {
                        const ExpTileTerm T = terms[q0 + j];
                        const double gs = (__popcll(base & T.zyo) & 1) ? -T.g : T.g;
                        zl[j] = T.zyl;
                        gr[j] = T.im ? 0.0 : gs;
                        gi[j] = T.im ? gs : 0.0;
                    }
                }
                auto pairs = [&](auto nqc) {
                    constexpr int NQ = decltype(nqc)::value;
                    // every thread has PPT pairs when the tile holds at least
                    // one per thread: no bound test, so the unrolled steps'
                    // LDS reads are all in flight together
                    constexpr bool FULL = TS / 2 >= (uint32_t)kExpThreads;
#pragma unroll
                    for (int u = 0; u < PPT; ++u) {
                        const uint32_t k = tid + (uint32_t)u * kExpThreads;
                        if (!FULL && k >= TS / 2) break;
                        const uint32_t e = ins0(k, G.hl);
                        const uint32_t se = swz(e);
                        const V a = tile[se], b = tile[se ^ sxl];  // swz is linear over XOR
                        const double zr = (double)a.x * (double)b.x + (double)a.y * (double)b.y;
                        const double zi = (double)a.x * (double)b.y - (double)a.y * (double)b.x;
                        double wr = 0.0, wi = 0.0;
#pragma unroll
                        for (int j = 0; j < NQ; ++j) {
                            const double sg = (__popc(e & zl[j]) & 1) ? -1.0 : 1.0;
                            wr += sg * gr[j];
                            wi += sg * gi[j];
                        }
                        acc += zr * wr + zi * wi;
                    }
                };
                switch (nq) {
                    case 1: pairs(std::integral_constant<int, 1>{}); break;
                    case 2: pairs(std::integral_constant<int, 2>{}); break;
                    case 3: pairs(std::integral_constant<int, 3>{}); break;
                    default: pairs(std::integral_constant<int, 4>{}); break;
                }
            }
        }
        __syncthreads();
    }
    double r = block_sum_e<kExpThreads>(acc, sh);
    if (tid == 0) partial[(uint64_t)blockIdx.y * gridDim.x + blockIdx.x] = r;
}

static int exp_tile_bits(int dtype) { return dtype == LQ_C128 ? 11 : 12; }

// LQ_EXPECT_TILES_AUTO: tiles for states over 64 MiB.  Smaller states stay
// in the Infinity Cache between the per-group reads, which are then cheaper
// than the tile kernel's LDS work.
static int g_exp_tiles = LQ_EXPECT_TILES_AUTO;

bool expect_tiles_wanted(int n, int dtype) {
    if (g_exp_tiles == LQ_EXPECT_TILES_OFF) return false;
    if (g_exp_tiles == LQ_EXPECT_TILES_ON) return true;
    const uint64_t bytes = (dtype == LQ_C128 ? 16ull : 8ull) << n;
    return bytes > (64ull << 20);
}

bool expect_tiles_batch_wanted(int n, int dtype, uint64_t ninst) {
    if (g_exp_tiles == LQ_EXPECT_TILES_OFF) return false;
    if (g_exp_tiles == LQ_EXPECT_TILES_ON) return true;
    return (double)ninst * std::ldexp(dtype == LQ_C128 ? 16.0 : 8.0, n) > (double)(256ull << 20);
}

void plan_expect_tiles(int n, int dtype, const PauliSum& ps, std::vector<ExpLayout>& layouts,
                       std::vector<int>& rest) {
    layouts.clear();
    rest.clear();
    const int K = std::min(exp_tile_bits(dtype), n);
    std::vector<int> flips;
    for (int g = 0; g < (int)ps.groups.size(); ++g)
        if (ps.groups[g].x != 0) flips.push_back(g);
    if (n < 8 || !expect_tiles_wanted(n, dtype)) {  // tiny or cache-resident states: the pair kernel
        rest = flips;
        return;
    }
    const int L = std::min(5, K);
    const int HB = K - L;
    const uint64_t low = (1ull << L) - 1ull;
    std::vector<int> todo;
    for (int g : flips) {
        if (__builtin_popcountll(ps.groups[g].x & ~low) <= HB) todo.push_back(g);
        else rest.push_back(g);
    }
    while (!todo.empty()) {
        uint64_t H = 0;
        std::vector<char> taken(todo.size(), 0);
        for (;;) {  // take the group needing the fewest new high bits
            int best = -1, bestc = 64;
            for (size_t i = 0; i < todo.size(); ++i) {
                if (taken[i]) continue;
                const uint64_t need = ps.groups[todo[i]].x & ~low & ~H;
                const int c = __builtin_popcountll(need);
                if (__builtin_popcountll(H | need) <= HB && c < bestc) {
                    best = (int)i;
                    bestc = c;
                }
            }
            if (best < 0) break;
            H |= ps.groups[todo[best]].x & ~low;
            taken[best] = 1;
        }
        for (int b = L; b < n && __builtin_popcountll(H) < HB; ++b) H |= 1ull << b;
        ExpLayout lay;
        std::memset(&lay.geo, 0, sizeof(lay.geo));
        lay.geo.L = L;
        for (int b = L; b < n; ++b)
            if (H & (1ull << b)) lay.geo.hpos[lay.geo.hb++] = b;
        lay.geo.ntiles = 1ull << (n - K);
        const uint64_t tilebits = low | H;
        auto local = [&](uint64_t m) -> uint32_t {
            uint32_t r = (uint32_t)(m & low);
            for (int j = 0; j < lay.geo.hb; ++j)
                if (m & (1ull << lay.geo.hpos[j])) r |= 1u << (L + j);
            return r;
        };
        std::vector<int> left;
        for (size_t i = 0; i < todo.size(); ++i) {
            const PauliGroup& grp = ps.groups[todo[i]];
            if ((grp.x & ~tilebits) != 0) {
                left.push_back(todo[i]);
                continue;
            }
            ExpTileGroup G;
            G.xl = local(grp.x);
            G.hl = 31 - __builtin_clz(G.xl);
            G.first = (int)lay.terms.size();
            G.count = grp.count;
            for (int t = grp.first; t < grp.first + grp.count; ++t) {
                const PauliTermDev& pt = ps.terms[t];
                ExpTileTerm T;
                T.zyo = pt.zy & ~tilebits;
                T.zyl = local(pt.zy & tilebits);
                T.im = __builtin_popcountll(grp.x & pt.zy) & 1;
                T.g = T.im ? -2.0 * pt.ci : 2.0 * pt.cr;
                lay.terms.push_back(T);
            }
            lay.groups.push_back(G);
        }
        layouts.push_back(std::move(lay));
        todo.swap(left);
    }
}

size_t expect_tile_lds(int dtype, const ExpGeom& geo) {
    const size_t amp = dtype == LQ_C128 ? 16 : 8;
    return std::max<size_t>((amp << (geo.L + geo.hb)) + (sizeof(uint64_t) << geo.hb), kExpThreads * sizeof(double));
}

int launch_expect_tile(const void* psi, int dtype, const ExpGeom& geo, uint64_t ninst, uint64_t inst_stride,
                       const ExpTileGroup* d_groups, int ngroups, const ExpTileTerm* d_terms, unsigned gx,
                       double* d_partial) {
    const size_t lds = expect_tile_lds(dtype, geo);
    const int K = geo.L + geo.hb;
    auto go = [&](auto kc, auto vt) {
        constexpr int KK = decltype(kc)::value;
        using V = decltype(vt);
        static bool attr = false;
        if (!attr) {
            (void)hipFuncSetAttribute((const void*)k_expect_tile<KK, V>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      96 * 1024);
            attr = true;
        }
        hipLaunchKernelGGL((k_expect_tile<KK, V>), dim3(gx, (unsigned)ninst), dim3(kExpThreads), lds, stream(),
                           (const V*)psi, geo, inst_stride, d_groups, ngroups, d_terms, d_partial);
    };
    auto by_k = [&](auto vt) {
        switch (K) {
            case 8: go(std::integral_constant<int, 8>{}, vt); break;
            case 9: go(std::integral_constant<int, 9>{}, vt); break;
            case 10: go(std::integral_constant<int, 10>{}, vt); break;
            case 11: go(std::integral_constant<int, 11>{}, vt); break;
            case 12: go(std::integral_constant<int, 12>{}, vt); break;
            case 13: go(std::integral_constant<int, 13>{}, vt); break;
            default: return false;
        }
        return true;
    };
    const bool ok = dtype == LQ_C128 ? by_k(double2{}) : by_k(float2{});
    if (!ok) {
        set_error("expectation tile of 2^%d amplitudes is not supported", K);
        return LQ_ERR_UNSUPPORTED;
    }
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

unsigned expect_tile_blocks(const ExpGeom& geo, uint64_t ninst) {
    const uint64_t want = std::max<uint64_t>(1, 2048 / std::max<uint64_t>(ninst, 1));
    return (unsigned)std::min<uint64_t>(geo.ntiles, want);
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

int orbit_max_groups(int dtype) {
    (void)dtype;
    return kOrbitMax;
}

// An item's terms are listed generator by generator (0 .. m-1), then the
// diagonal terms: the orbit kernel walks them in that order.
void plan_orbits(const PauliSum& ps, const std::vector<int>& groups, int diag, int maxm, std::vector<OrbitSet>& out) {
    out.clear();
    maxm = std::max(1, std::min(maxm, kOrbitMax));
    // per set: a reduced basis (each vector has exactly one of the set's
    // pivot bits), its pivots, and the member groups (the generators)
    struct Build {
        std::vector<uint64_t> basis;
        std::vector<int> pivb, gens;
    };
    std::vector<Build> builds;
    for (int g : groups) {
        const uint64_t x = ps.groups[g].x;
        bool placed = false;
        for (Build& b : builds) {
            if ((int)b.gens.size() >= maxm) continue;
            uint64_t r = x;
            for (size_t k = 0; k < b.basis.size(); ++k)
                if ((r >> b.pivb[k]) & 1ull) r ^= b.basis[k];
            if (!r) continue;  // x lies in the set's span
            const int p = 63 - __builtin_clzll(r);
            for (uint64_t& v : b.basis)
                if ((v >> p) & 1ull) v ^= r;
            b.basis.push_back(r);
            b.pivb.push_back(p);
            b.gens.push_back(g);
            placed = true;
            break;
        }
        if (!placed) {
            Build b;
            b.basis.push_back(x);
            b.pivb.push_back(63 - __builtin_clzll(x));
            b.gens.push_back(g);
            builds.push_back(b);
        }
    }
    auto sigma = [](const OrbitSet& o, uint64_t zy) {
        uint32_t s = 0;
        for (int c = 0; c < (1 << o.m); ++c) s |= (uint32_t)(__builtin_popcountll(o.X[c] & zy) & 1) << c;
        return s;
    };
    for (const Build& b : builds) {
        OrbitSet o;
        o.m = (int)b.gens.size();
        std::vector<int> pv = b.pivb;
        std::sort(pv.begin(), pv.end());
        for (int j = 0; j < o.m; ++j) o.piv[j] = pv[j];
        for (int c = 0; c < (1 << o.m); ++c) {
            uint64_t xc = 0;
            for (int j = 0; j < o.m; ++j)
                if ((c >> j) & 1) xc ^= ps.groups[b.gens[j]].x;
            o.X[c] = xc;
        }
        for (int j = 0; j < o.m; ++j) {
            const PauliGroup& grp = ps.groups[b.gens[j]];
            for (int t = grp.first; t < grp.first + grp.count; ++t) {
                const PauliTermDev& d = ps.terms[t];
                OrbitTerm q;
                std::memset(&q, 0, sizeof(q));
                q.zy = d.zy;
                q.gen = j;
                q.imag = __builtin_popcountll(grp.x & d.zy) & 1;
                q.g = q.imag ? -2.0 * d.ci : 2.0 * d.cr;
                q.sigma = sigma(o, d.zy);
                o.terms.push_back(q);
            }
        }
        out.push_back(o);
    }
    if (diag >= 0 && ps.groups[diag].count > 0) {
        if (out.empty()) out.emplace_back();
        OrbitSet& o = out[0];
        const PauliGroup& grp = ps.groups[diag];
        for (int t = grp.first; t < grp.first + grp.count; ++t) {
            OrbitTerm q;
            std::memset(&q, 0, sizeof(q));
            q.zy = ps.terms[t].zy;
            q.gen = -1;
            q.g = ps.terms[t].cr;
            q.sigma = sigma(o, q.zy);
            o.terms.push_back(q);
        }
    }
}

// Device memory an expectation works in when the caller supplies it (the
// no-allocation entry point): its base and size.  Null: the library's
// growable scratch and pinned staging.
struct ExpWorkspace {
    void* base;
    size_t bytes;
};

template <typename V>
static int expect_t(const lq_state* st, const PauliSum& ps, double* out, const ExpWorkspace* ws) {
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
    // Device term table: the diagonal group's terms are reordered static
    // first (masks that miss the r bits), then dynamic; a launch takes every
    // static term and up to kDiagDyn dynamic ones, further launches the rest.
    std::vector<PauliTermDev> dev = ps.terms;
    int diag_first = -1, diag_static = 0, diag_count = 0;
    size_t launches = 0;
    // flip groups: from LDS tiles where they fit a layout, else orbit items
    // (several groups per read of the state); the run-time specialised
    // layouts when the policy allows and all compile
    std::vector<ExpLayout> lays;
    std::vector<int> rest;
    XPlan xp;
    // (not for a caller-supplied workspace: compiling and loading a kernel allocate)
    const bool use_x = !ws && expect_jit_wanted(st->n, st->dtype) && plan_expect_jit(st->n, st->dtype, ps, xp) &&
                       prepare_expect_jit(xp, false, true);
    int x_gx = 0;
    if (use_x) {
        rest = xp.rest;
        x_gx = expect_jit_blocks(xp, 1);
    } else {
        plan_expect_tiles(st->n, st->dtype, ps, lays, rest);
    }
    const bool diag_in_tiles = use_x && xp.diag_in;
    std::vector<char> paired(ng, 0);
    for (int g : rest) paired[g] = 1;
    size_t tile_parts = use_x ? (size_t)x_gx * xp.layouts.size() : 0;
    std::vector<unsigned> lay_blocks;
    for (const ExpLayout& l : lays) {
        lay_blocks.push_back(expect_tile_blocks(l.geo, 1));
        tile_parts += lay_blocks.back();
    }
    std::vector<int> pair_groups;
    for (int g = 0; g < ng; ++g)
        if (paired[g] && ps.groups[g].x != 0) pair_groups.push_back(g);
    std::vector<OrbitSet> osets;
    plan_orbits(ps, pair_groups, -1, orbit_max_groups(st->dtype), osets);
    std::vector<OrbitItem> oitems;
    size_t oterms = 0, ox = 0;
    int oblocks = 0;
    for (const OrbitSet& o : osets) {
        OrbitItem it;
        std::memset(&it, 0, sizeof(it));
        it.work = st->dim >> o.m;
        it.m = o.m;
        for (int q = 0; q < o.m; ++q) it.piv[q] = o.piv[q];
        it.x_first = (int)ox;
        it.term_first = (int)oterms;
        it.term_count = (int)o.terms.size();
        it.block_first = oblocks;
        it.nblocks = (int)std::min<uint64_t>(std::max<uint64_t>((it.work + 1023) / 1024, 1), kExpBlocksMax);
        oblocks += it.nblocks;
        oterms += o.terms.size();
        ox += (size_t)1 << o.m;
        oitems.push_back(it);
    }
    for (const PauliGroup& grp : ps.groups) {
        if (grp.x != 0 || diag_in_tiles) continue;
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
    // layout tables after the term table: [groups][terms] per layout
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    std::vector<size_t> lay_off;
    size_t lay_bytes = 0;
    for (const ExpLayout& l : lays) {
        lay_off.push_back(lay_bytes);
        lay_bytes += al(l.groups.size() * sizeof(ExpTileGroup)) + al(l.terms.size() * sizeof(ExpTileTerm));
    }
    const size_t x_coef_off = lay_bytes;  // the specialised layouts' coefficient table
    if (use_x) lay_bytes += al(expect_jit_coef_count(xp) * sizeof(double));
    // orbit tables: [items][terms][X]
    const size_t o_items_off = lay_bytes;
    lay_bytes += al(oitems.size() * sizeof(OrbitItem));
    const size_t o_terms_off = lay_bytes;
    lay_bytes += al(oterms * sizeof(OrbitTerm));
    const size_t o_x_off = lay_bytes;
    lay_bytes += al(ox * sizeof(uint64_t));
    size_t part_count = launches * nb_diag + tile_parts + oblocks;
    const size_t need = term_pad + lay_bytes + (part_count + 8) * sizeof(double);
    char* base;
    double* h;
    std::vector<double> pageable;  // the tables' host copy when no pinned staging may be allocated
    if (ws) {
        if (need > ws->bytes) {
            set_error("this Pauli sum needs %zu bytes of workspace on a %d-qubit state; the static workspace holds %zu",
                      need, st->n, ws->bytes);
            return LQ_ERR_INVALID_ARGUMENT;
        }
        base = (char*)ws->base;
        pageable.resize((term_pad + lay_bytes) / sizeof(double) + 16);
        h = pageable.data();
    } else {
        base = (char*)scratch(need);
        if (!base) {
            set_error("scratch allocation failed");
            return LQ_ERR_OUT_OF_MEMORY;
        }
        h = host_staging((term_pad + lay_bytes) / sizeof(double) + 16);
        if (!h) return LQ_ERR_OUT_OF_MEMORY;
    }
    PauliTermDev* d_terms = (PauliTermDev*)base;
    char* d_lay = base + term_pad;
    double* d_part = (double*)(base + term_pad + lay_bytes);
    double* d_out = d_part + part_count;
    std::memcpy(h, dev.data(), term_bytes);
    if (use_x) expect_jit_coefs(xp, (double*)((char*)h + term_pad + x_coef_off));
    if (!oitems.empty()) {
        char* ho = (char*)h + term_pad;
        std::memcpy(ho + o_items_off, oitems.data(), oitems.size() * sizeof(OrbitItem));
        size_t tb = 0, xb = 0;
        for (const OrbitSet& o : osets) {
            std::memcpy(ho + o_terms_off + tb * sizeof(OrbitTerm), o.terms.data(), o.terms.size() * sizeof(OrbitTerm));
            std::memcpy(ho + o_x_off + xb * sizeof(uint64_t), o.X, ((size_t)1 << o.m) * sizeof(uint64_t));
            tb += o.terms.size();
            xb += (size_t)1 << o.m;
        }
    }
    for (size_t li = 0; li < lays.size(); ++li) {
        char* hp = (char*)h + term_pad + lay_off[li];
        const ExpLayout& l = lays[li];
        std::memcpy(hp, l.groups.data(), l.groups.size() * sizeof(ExpTileGroup));
        std::memcpy(hp + al(l.groups.size() * sizeof(ExpTileGroup)), l.terms.data(),
                    l.terms.size() * sizeof(ExpTileTerm));
    }
    LQ_HIP_CHECK(hipMemcpyAsync(d_terms, h, term_pad + lay_bytes, hipMemcpyHostToDevice, stream()));
    int written = 0;
    if (use_x) {
        int rc = launch_expect_jit(xp, st->d_amps, (const double*)(d_lay + x_coef_off), 0, 1, 0, x_gx,
                                   d_part + written);
        if (rc) return rc;
        written += x_gx * (int)xp.layouts.size();
    }
    for (size_t li = 0; li < lays.size(); ++li) {
        const ExpLayout& l = lays[li];
        const char* dp = d_lay + lay_off[li];
        int rc = launch_expect_tile(st->d_amps, st->dtype, l.geo, 1, 0, (const ExpTileGroup*)dp,
                                    (int)l.groups.size(),
                                    (const ExpTileTerm*)(dp + al(l.groups.size() * sizeof(ExpTileGroup))),
                                    lay_blocks[li], d_part + written);
        if (rc) return rc;
        written += (int)lay_blocks[li];
    }
    for (int g = 0; g < ng; ++g) {
        const PauliGroup& grp = ps.groups[g];
        if (grp.x == 0 && diag_in_tiles) continue;
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
        }
        LQ_LAUNCH_CHECK();
    }
    if (oblocks > 0) {
        hipLaunchKernelGGL(k_expect_orbits<V>, dim3(oblocks), dim3(kOrbitThreads), 0, stream(), (const V*)st->d_amps,
                           (const OrbitItem*)(d_lay + o_items_off), (int)oitems.size(),
                           (const uint64_t*)(d_lay + o_x_off), (const OrbitTerm*)(d_lay + o_terms_off),
                           d_part + written);
        LQ_LAUNCH_CHECK();
        written += oblocks;
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

static int expect_ws(const lq_state* st, const PauliSum& ps, double* out, const ExpWorkspace* ws) {
    if (ps.n != st->n) {
        // "Pauli term must have operator for each qubit" (observable.rs:97-101)
        set_error("Pauli term must have operator for each qubit (%d terms' qubits vs state %d)", ps.n, st->n);
        return LQ_ERR_DIMENSION_MISMATCH;
    }
    if (st->dtype == LQ_C128) return expect_t<double2>(st, ps, out, ws);
    return expect_t<float2>(st, ps, out, ws);
}

int expect_pauli_sum(const lq_state* st, const PauliSum& ps, double* out) { return expect_ws(st, ps, out, nullptr); }

}  // namespace lq

// Workspace of lq_expect_pauli_sum_noalloc: term tables, tile layouts and
// per-block partials.  Every GPU the library runs on holds its own copy until
// the process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (g_noalloc_ptr), never
// through this name.
static constexpr size_t kNoAllocBytes = 4u << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kNoAllocDevices = 64;
void* g_noalloc_ptr[kNoAllocDevices] = {};

// The workspace address on the current device, looked up on the first call
// made on that device (under the library lock).
int noalloc_workspace(void** out) {
    int dev = 0;
    LQ_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) {
        lq::set_error("device %d is beyond the %d devices the static workspace table holds", dev, kNoAllocDevices);
        return LQ_ERR_UNSUPPORTED;
    }
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
        if (e != hipSuccess || !p) {
            lq::set_error("hipGetSymbolAddress of the static workspace failed: %s", hipGetErrorString(e));
            return LQ_ERR_DEVICE;
        }
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return LQ_OK;
}
}  // namespace

using namespace lq;

extern "C" int lq_expect_plan(int num_qubits, int dtype, const uint8_t* paulis, const double* coeffs, int num_terms,
                              int convention, int* tile_layouts, int* pair_groups, int* diag_groups) {
    LQ_API_GUARD;
    clear_error();
    if (num_qubits < 1 || num_qubits > kMaxQubits || num_terms < 0 || (num_terms > 0 && (!paulis || !coeffs)) ||
        (dtype != LQ_C128 && dtype != LQ_C64)) {
        set_error("invalid arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    PauliSum ps;
    int rc = compile_pauli_sum(num_qubits, paulis, coeffs, num_terms, convention, &ps);
    if (rc) return rc;
    int diag = 0;
    for (const PauliGroup& g : ps.groups) diag += g.x == 0;
    // reads for the flip groups outside the tiles: one per orbit item
    auto orbit_reads = [&](const std::vector<int>& rest) {
        std::vector<int> flip;
        for (int g : rest)
            if (ps.groups[g].x != 0) flip.push_back(g);
        std::vector<OrbitSet> sets;
        plan_orbits(ps, flip, -1, orbit_max_groups(dtype), sets);
        return (int)sets.size();
    };
    XPlan xp;
    if (expect_jit_wanted(num_qubits, dtype) && plan_expect_jit(num_qubits, dtype, ps, xp)) {
        // the run-time specialised layouts (used once their kernels compile)
        if (tile_layouts) *tile_layouts = (int)xp.layouts.size();
        if (pair_groups) *pair_groups = orbit_reads(xp.rest);
        if (diag_groups) *diag_groups = xp.diag_in ? 0 : diag;
        return LQ_OK;
    }
    std::vector<ExpLayout> lays;
    std::vector<int> rest;
    plan_expect_tiles(num_qubits, dtype, ps, lays, rest);
    if (tile_layouts) *tile_layouts = (int)lays.size();
    if (pair_groups) *pair_groups = orbit_reads(rest);
    if (diag_groups) *diag_groups = diag;
    return LQ_OK;
}

extern "C" int lq_expect_jit_compile(int num_qubits, int dtype, const uint8_t* paulis, const double* coeffs,
                                     int num_terms, int convention, int* kernels) {
    LQ_API_GUARD;
    clear_error();
    if (num_qubits < 1 || num_qubits > kMaxQubits || num_terms < 0 || (num_terms > 0 && (!paulis || !coeffs)) ||
        (dtype != LQ_C128 && dtype != LQ_C64)) {
        set_error("invalid arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    PauliSum ps;
    int rc = compile_pauli_sum(num_qubits, paulis, coeffs, num_terms, convention, &ps);
    if (rc) return rc;
    if (kernels) *kernels = 0;
    XPlan xp;
    if (!expect_jit_wanted(num_qubits, dtype) || !plan_expect_jit(num_qubits, dtype, ps, xp)) return LQ_OK;
    if (!prepare_expect_jit(xp, true, false)) {
        set_error("hipRTC failed on an expectation kernel (LQ_JIT_LOG=1 prints the log)");
        return LQ_ERR_DEVICE;
    }
    if (kernels) *kernels = (int)xp.layouts.size();
    return LQ_OK;
}

extern "C" int lq_set_expect_tiles(int mode) {
    LQ_API_GUARD;
    clear_error();
    if (mode != LQ_EXPECT_TILES_OFF && mode != LQ_EXPECT_TILES_AUTO && mode != LQ_EXPECT_TILES_ON) {
        set_error("mode must be LQ_EXPECT_TILES_OFF, LQ_EXPECT_TILES_AUTO or LQ_EXPECT_TILES_ON");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    g_exp_tiles = mode;
    return LQ_OK;
}

extern "C" int lq_expect_pauli_sum(const lq_state* st, const uint8_t* paulis, const double* coeffs, int num_terms,
                                   int convention, double* out) {
    LQ_API_GUARD;
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

extern "C" int lq_expect_pauli_sum_noalloc(const lq_state* st, const uint8_t* paulis, const double* coeffs,
                                           int num_terms, int convention, double* out) {
    LQ_API_GUARD;
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
    ExpWorkspace ws{nullptr, kNoAllocBytes};
    rc = noalloc_workspace(&ws.base);
    if (rc) return rc;
    return expect_ws(st, ps, out, &ws);
}

extern "C" int lq_expect_jit_describe(int num_qubits, int dtype, const uint8_t* paulis, const double* coeffs,
                                      int num_terms, int convention, char* out, size_t out_len, size_t* needed) {
    LQ_API_GUARD;
    clear_error();
    if (num_qubits < 1 || num_qubits > kMaxQubits || num_terms < 0 || (num_terms > 0 && (!paulis || !coeffs)) ||
        (dtype != LQ_C128 && dtype != LQ_C64)) {
        set_error("invalid arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    PauliSum ps;
    int rc = compile_pauli_sum(num_qubits, paulis, coeffs, num_terms, convention, &ps);
    if (rc) return rc;
    XPlan xp;
    std::string js = "{}";
    if (expect_jit_wanted(num_qubits, dtype) && plan_expect_jit(num_qubits, dtype, ps, xp)) js = describe_expect_jit(xp);
    if (needed) *needed = js.size() + 1;
    if (out && out_len > 0) {
        const size_t k = std::min(out_len - 1, js.size());
        std::memcpy(out, js.data(), k);
        out[k] = 0;
    }
    return LQ_OK;
}
