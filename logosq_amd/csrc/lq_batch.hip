// Batched execution: many states of one size run the same pass plan with
// per-instance op tables (their own angles), one launch per pass for the
// whole batch; their Pauli-sum expectations are reduced by one work-item
// kernel over every (instance, flip-mask group).  Used for
//   - all 2P+1 parameter-shift circuits of a gradient (the reference runs them
//     one after another: src/optimization/gradient.rs:53-74), and
//   - batches of unrelated circuits of uneven qubit counts.
// Accumulation is f64 for both amplitude types, in a fixed order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "lq_fused.h"

namespace lq {

void bind_template(const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                   std::vector<lq_gate>& out);

// ---------------------------------------------------------------------------------
// device buffers reused across calls
// ---------------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    int get(size_t bytes) {
        if (bytes <= cap) return LQ_OK;
        LQ_HIP_CHECK(hipStreamSynchronize(stream()));
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
            return LQ_ERR_OUT_OF_MEMORY;
        }
        cap = bytes;
        return LQ_OK;
    }
};

struct HostBuf {
    void* p = nullptr;
    size_t cap = 0;
    int get(size_t bytes) {
        if (bytes <= cap) return LQ_OK;
        LQ_HIP_CHECK(hipStreamSynchronize(stream()));
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        LQ_HIP_CHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
        cap = bytes;
        return LQ_OK;
    }
};

static DevBuf g_states, g_bops, g_items, g_terms, g_part, g_out;
static HostBuf g_hops, g_hitems, g_hterms, g_hout;

// ---------------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------------
template <typename V>
__global__ void k_fill_zero_states(V* __restrict__ psi, uint64_t dim, uint64_t total) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        V x;
        x.x = (i & (dim - 1)) == 0 ? 1 : 0;
        x.y = 0;
        psi[i] = x;
    }
}

// One op whose matrix differs from instance 0's in instance `inst`.
struct OpOverride {
    int32_t inst, op;
    double m[32];
};
static_assert(sizeof(TileOp) % 8 == 0, "op tables are replicated in 8-byte words");

constexpr int kEThreads = 256;

template <int NT>
__device__ __forceinline__ double bsum(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    return sh[0];
}

// Every instance's op table becomes a copy of instance 0's (8-byte words).
__global__ void __launch_bounds__(256) k_replicate_ops(uint64_t* __restrict__ dst, const uint64_t* __restrict__ base,
                                                      uint64_t words, uint64_t ninst) {
    const uint64_t total = words * ninst;
    for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (uint64_t)gridDim.x * 256)
        dst[t] = base[t % words];
}

// One block per override: instance inst's op gets its own matrix.
__global__ void __launch_bounds__(64) k_apply_overrides(TileOp* __restrict__ ops, const OpOverride* __restrict__ ov,
                                                       uint64_t nops) {
    const OpOverride& o = ov[blockIdx.x];
    if (threadIdx.x < 32) ops[(uint64_t)o.inst * nops + o.op].m[threadIdx.x] = o.m[threadIdx.x];
}

// States first .. first+count-1 of the batch become copies of state 0.
template <typename V>
__global__ void __launch_bounds__(256) k_copy_base(V* __restrict__ psi, uint64_t dim, uint64_t first, uint64_t count) {
    const uint64_t total = dim * count;
    for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (uint64_t)gridDim.x * 256) {
        const uint64_t c = t / dim, i = t - c * dim;
        psi[(first + c) * dim + i] = psi[i];
    }
}

// energies[inst] = sum of the partials of the instance's blocks, then of
// its LDS-tile partials (tpart[layout][inst][gx]), in a fixed order.
__global__ void __launch_bounds__(256) k_sum_ranges(const double* __restrict__ partial,
                                                   const int* __restrict__ first, const int* __restrict__ count,
                                                   const double* __restrict__ tpart, int nl, int gx, int ninst,
                                                   double* __restrict__ out) {
    __shared__ double sh[256];
    const int inst = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < count[inst]; i += 256) s += partial[first[inst] + i];
    for (int q = threadIdx.x; q < nl * gx; q += 256)
        s += tpart[((size_t)(q / gx) * ninst + inst) * gx + q % gx];
    double t = bsum<256>(s, sh);
    if (threadIdx.x == 0) out[inst] = t;
}

// ---------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------
// Per-instance Pauli sums: term tables concatenated, groups per instance.
struct BatchObs {
    const PauliSum* ps;  // instance's compiled sum
    int term_base;       // offset of its terms in the concatenated table
};

static int blocks_for(uint64_t work) {
    uint64_t b = (work + kEThreads * 8 - 1) / (kEThreads * 8);
    if (b < 1) b = 1;
    if (b > 1024) b = 1024;
    return (int)b;
}

// One batch: ninst states of dim amplitudes run one pass plan (per-instance
// op tables) from |0...0>, then every instance's Pauli-sum expectation.  The
// host side prepares everything into pinned buffers first; the device work
// is then one launch sequence on the library stream (fill, op upload, tile
// passes, item upload, expectation, sums, result copy), which can be
// captured once into a hipGraph and replayed for later batches of the same
// shape: only buffer contents change between replays.
static DevBuf g_bplan;
static HostBuf g_hplan;

struct BatchJob {
    int n = 0, dtype = 0, ninst = 0;
    uint64_t dim = 0;
    const Plan* plan = nullptr;
    size_t nops = 0, goff = 0, plan_bytes = 0;
    int nitems = 0, blocks = 0;
    size_t roff = 0, items_bytes = 0, terms_bytes = 0, xoff = 0;
    // flip groups evaluated from LDS tiles (when every instance shares one
    // sum): layouts, their tables in the items buffer, partials after the
    // items' partials
    std::vector<ExpLayout> lays;
    std::vector<size_t> lay_off;
    int tile_gx = 0;
    size_t tpart_off = 0;
    // shared prefixes: instances are ordered by the step where they leave
    // instance 0's circuit; alive[k] instances take part in step k (the rest
    // still equal instance 0) and alive[steps] = ninst.  Empty: all, always.
    std::vector<int> alive;
    // compact op tables: the host holds instance 0's table and per-instance
    // overrides (hbase bytes, then nover records); the device replicates the
    // table into every instance's slot and applies the overrides
    bool replicate = false;
    size_t nover = 0, hgoff = 0, hover = 0, hbytes = 0, stage = 0;
};


// Compact op tables: pinned [instance 0's ops][groups][nover overrides]
// (filled by the caller: *base and *over); the device holds every
// instance's table, the groups, then a staging copy of the host layout.
static int job_ops_compact(BatchJob& j, size_t nover, TileOp** base, OpOverride** over) {
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t ob = j.nops * (size_t)j.ninst * sizeof(TileOp);
    j.replicate = true;
    j.nover = nover;
    j.goff = al(ob);
    const size_t gb = j.plan->groups.size() * sizeof(TileGroup);
    j.plan_bytes = j.goff + gb;
    j.hgoff = al(j.nops * sizeof(TileOp));
    j.hover = al(j.hgoff + gb);
    j.hbytes = j.hover + nover * sizeof(OpOverride);
    j.stage = al(j.plan_bytes);
    int rc = g_hplan.get(j.hbytes + 256);
    if (!rc) rc = g_bplan.get(j.stage + j.hbytes + 256);
    if (rc) return rc;
    if (gb) std::memcpy((char*)g_hplan.p + j.hgoff, j.plan->groups.data(), gb);
    *base = (TileOp*)g_hplan.p;
    *over = (OpOverride*)((char*)g_hplan.p + j.hover);
    return LQ_OK;
}

// Pinned space for the op tables of ninst instances (filled by the caller),
// followed by the plan's groups.
static int job_ops(BatchJob& j, TileOp** out) {
    const size_t ob = j.nops * (size_t)j.ninst * sizeof(TileOp);
    j.goff = (ob + 255) & ~(size_t)255;
    j.plan_bytes = j.goff + j.plan->groups.size() * sizeof(TileGroup);
    int rc = g_hplan.get(j.plan_bytes + 256);
    if (!rc) rc = g_bplan.get(j.plan_bytes + 256);
    if (rc) return rc;
    if (!j.plan->groups.empty())
        std::memcpy((char*)g_hplan.p + j.goff, j.plan->groups.data(), j.plan->groups.size() * sizeof(TileGroup));
    *out = (TileOp*)g_hplan.p;
    return LQ_OK;
}

// Expectation work of the batch (obs_of[i] indexes `sums`): with one sum
// shared by every instance, the flip groups that fit an LDS tile layout are
// read from tiles; the other flip groups go into orbit items (up to
// orbit_max_groups groups per read of a state) with each instance's diagonal
// terms in its first item.  Items, per-instance block ranges and layout
// tables go into the items buffer; orbit terms, then the X tables, into the
// terms buffer.
static int job_expect(BatchJob& j, const std::vector<PauliSum>& sums, const std::vector<int>& obs_of) {
    std::vector<OrbitItem> items;
    std::vector<int> first(j.ninst), count(j.ninst);
    int blocks = 0;
    j.lays.clear();
    j.lay_off.clear();
    j.tile_gx = 0;
    std::vector<char> tiled;
    if (sums.size() == 1) {
        std::vector<int> rest;
        plan_expect_tiles(j.n, j.dtype, sums[0], j.lays, rest);
        tiled.assign(sums[0].groups.size(), 0);
        for (size_t g = 0; g < tiled.size(); ++g) tiled[g] = sums[0].groups[g].x != 0;
        for (int g : rest) tiled[g] = 0;
        if (!j.lays.empty()) j.tile_gx = (int)expect_tile_blocks(j.lays[0].geo, (uint64_t)j.ninst);
    }
    // orbit sets of each distinct sum, their terms and X tables concatenated
    const int maxm = orbit_max_groups(j.dtype);
    std::vector<std::vector<OrbitSet>> sets(sums.size());
    std::vector<int> term_base(sums.size()), x_base(sums.size());
    size_t nterms = 0, nx = 0;
    for (size_t k = 0; k < sums.size(); ++k) {
        std::vector<int> flip;
        int diag = -1;
        for (size_t g = 0; g < sums[k].groups.size(); ++g) {
            if (sums[k].groups[g].x == 0) diag = (int)g;
            else if (tiled.empty() || !tiled[g]) flip.push_back((int)g);
        }
        plan_orbits(sums[k], flip, diag, maxm, sets[k]);
        term_base[k] = (int)nterms;
        x_base[k] = (int)nx;
        for (const OrbitSet& o : sets[k]) {
            nterms += o.terms.size();
            nx += (size_t)1 << o.m;
        }
    }
    for (int i = 0; i < j.ninst; ++i) {
        const int k = obs_of[i];
        first[i] = blocks;
        int tb = term_base[k], xb = x_base[k];
        for (const OrbitSet& o : sets[k]) {
            OrbitItem it;
            std::memset(&it, 0, sizeof(it));
            it.state_off = (uint64_t)i * j.dim;
            it.work = j.dim >> o.m;
            it.m = o.m;
            for (int q = 0; q < o.m; ++q) it.piv[q] = o.piv[q];
            it.x_first = xb;
            it.term_first = tb;
            it.term_count = (int)o.terms.size();
            it.block_first = blocks;
            it.nblocks = blocks_for(it.work);
            it.inst = i;
            blocks += it.nblocks;
            tb += (int)o.terms.size();
            xb += 1 << o.m;
            items.push_back(it);
        }
        count[i] = blocks - first[i];
    }
    j.nitems = (int)items.size();
    j.blocks = blocks;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t ib = items.size() * sizeof(OrbitItem);
    const size_t rb = (size_t)j.ninst * 2 * sizeof(int);
    j.roff = al(ib);
    j.items_bytes = al(j.roff + rb);
    for (const ExpLayout& l : j.lays) {
        j.lay_off.push_back(j.items_bytes);
        j.items_bytes += al(l.groups.size() * sizeof(ExpTileGroup)) + al(l.terms.size() * sizeof(ExpTileTerm));
    }
    j.xoff = al(std::max<size_t>(nterms, 1) * sizeof(OrbitTerm));
    j.terms_bytes = j.xoff + std::max<size_t>(nx, 1) * sizeof(uint64_t);
    j.tpart_off = (size_t)std::max(blocks, 1);
    const size_t tparts = j.lays.size() * (size_t)j.ninst * (size_t)j.tile_gx;
    int rc = g_items.get(j.items_bytes + 256);
    if (!rc) rc = g_terms.get(j.terms_bytes);
    if (!rc) rc = g_part.get((j.tpart_off + tparts + 1) * sizeof(double));
    if (!rc) rc = g_out.get((size_t)j.ninst * sizeof(double));
    if (!rc) rc = g_hitems.get(j.items_bytes + 256);
    if (!rc) rc = g_hterms.get(j.terms_bytes);
    if (!rc) rc = g_hout.get((size_t)j.ninst * sizeof(double));
    if (rc) return rc;
    char* hi = (char*)g_hitems.p;
    if (ib) std::memcpy(hi, items.data(), ib);
    std::memcpy(hi + j.roff, first.data(), j.ninst * sizeof(int));
    std::memcpy(hi + j.roff + j.ninst * sizeof(int), count.data(), j.ninst * sizeof(int));
    for (size_t l = 0; l < j.lays.size(); ++l) {
        const ExpLayout& lay = j.lays[l];
        std::memcpy(hi + j.lay_off[l], lay.groups.data(), lay.groups.size() * sizeof(ExpTileGroup));
        std::memcpy(hi + j.lay_off[l] + al(lay.groups.size() * sizeof(ExpTileGroup)), lay.terms.data(),
                    lay.terms.size() * sizeof(ExpTileTerm));
    }
    OrbitTerm* ht = (OrbitTerm*)g_hterms.p;
    uint64_t* hx = (uint64_t*)((char*)g_hterms.p + j.xoff);
    for (size_t k = 0; k < sums.size(); ++k) {
        int tb = term_base[k], xb = x_base[k];
        for (const OrbitSet& o : sets[k]) {
            if (!o.terms.empty()) std::memcpy(ht + tb, o.terms.data(), o.terms.size() * sizeof(OrbitTerm));
            std::memcpy(hx + xb, o.X, ((size_t)1 << o.m) * sizeof(uint64_t));
            tb += (int)o.terms.size();
            xb += 1 << o.m;
        }
    }
    return LQ_OK;
}

// The device work of a prepared job (capturable when the plan has no
// renormalisation step).
static int copy_base(const BatchJob& j, void* psi, int first, int count) {
    if (count <= 0) return LQ_OK;
    const uint64_t total = j.dim * (uint64_t)count;
    const unsigned g = (unsigned)std::min<uint64_t>((total + 255) / 256, 65536);
    if (j.dtype == LQ_C128)
        hipLaunchKernelGGL(k_copy_base<double2>, dim3(g), dim3(256), 0, stream(), (double2*)psi, j.dim,
                           (uint64_t)first, (uint64_t)count);
    else
        hipLaunchKernelGGL(k_copy_base<float2>, dim3(g), dim3(256), 0, stream(), (float2*)psi, j.dim,
                           (uint64_t)first, (uint64_t)count);
    LQ_LAUNCH_CHECK();
    return LQ_OK;
}

static int job_launch(const BatchJob& j, void* psi) {
    const size_t nsteps = j.plan->steps.size();
    auto alive = [&](size_t k) { return j.alive.empty() ? j.ninst : j.alive[k]; };
    // instances that diverge in the first step start from |0...0> too
    const uint64_t total = j.dim * (uint64_t)alive(0);
    const unsigned g = (unsigned)std::min<uint64_t>((total + 255) / 256, 65536);
    if (j.dtype == LQ_C128)
        hipLaunchKernelGGL(k_fill_zero_states<double2>, dim3(g), dim3(256), 0, stream(), (double2*)psi, j.dim, total);
    else
        hipLaunchKernelGGL(k_fill_zero_states<float2>, dim3(g), dim3(256), 0, stream(), (float2*)psi, j.dim, total);
    LQ_LAUNCH_CHECK();
    if (j.replicate) {
        char* dev = (char*)g_bplan.p;
        LQ_HIP_CHECK(hipMemcpyAsync(dev + j.stage, g_hplan.p, j.hbytes, hipMemcpyHostToDevice, stream()));
        const size_t gb = j.plan->groups.size() * sizeof(TileGroup);
        if (gb)
            LQ_HIP_CHECK(hipMemcpyAsync(dev + j.goff, dev + j.stage + j.hgoff, gb, hipMemcpyDeviceToDevice, stream()));
        const uint64_t words = j.nops * sizeof(TileOp) / 8;
        const uint64_t tw = words * (uint64_t)j.ninst;
        const unsigned g2 = (unsigned)std::min<uint64_t>((tw + 255) / 256, 65536);
        hipLaunchKernelGGL(k_replicate_ops, dim3(g2), dim3(256), 0, stream(), (uint64_t*)dev,
                           (const uint64_t*)(dev + j.stage), words, (uint64_t)j.ninst);
        LQ_LAUNCH_CHECK();
        if (j.nover) {
            hipLaunchKernelGGL(k_apply_overrides, dim3((unsigned)j.nover), dim3(64), 0, stream(), (TileOp*)dev,
                               (const OpOverride*)(dev + j.stage + j.hover), (uint64_t)j.nops);
            LQ_LAUNCH_CHECK();
        }
    } else {
        LQ_HIP_CHECK(hipMemcpyAsync(g_bplan.p, g_hplan.p, j.plan_bytes, hipMemcpyHostToDevice, stream()));
    }
    const TileOp* d_ops = (const TileOp*)g_bplan.p;
    const TileGroup* d_groups = (const TileGroup*)((char*)g_bplan.p + j.goff);
    for (size_t k = 0; k < nsteps; ++k) {
        const Step& s = j.plan->steps[k];
        const int na = alive(k);
        int rc;
        if (k > 0 && na > alive(k - 1)) {
            // instances leaving instance 0's circuit at this step: its state so far
            rc = copy_base(j, psi, alive(k - 1), na - alive(k - 1));
            if (rc) return rc;
        }
        if (s.normalize) {
            // MatrixGate renormalisation: per instance
            for (int i = 0; i < na; ++i) {
                lq_state tmp;
                tmp.n = j.n;
                tmp.dtype = j.dtype;
                tmp.dim = j.dim;
                tmp.amp_bytes = j.dtype == LQ_C128 ? 16 : 8;
                tmp.d_amps = (char*)psi + (size_t)i * j.dim * tmp.amp_bytes;
                rc = normalize_device(&tmp, nullptr);
                if (rc) return rc;
            }
            continue;
        }
        rc = launch_tile_pass(psi, j.dtype, s.pass, d_ops, d_groups, (uint64_t)na, j.dim, j.nops, s.jit);
        if (rc) return rc;
    }
    if (nsteps > 0 && alive(nsteps - 1) < j.ninst) {
        int rc = copy_base(j, psi, alive(nsteps - 1), j.ninst - alive(nsteps - 1));
        if (rc) return rc;
    }
    if (j.blocks > 0 || !j.lays.empty()) {
        LQ_HIP_CHECK(hipMemcpyAsync(g_items.p, g_hitems.p, j.items_bytes, hipMemcpyHostToDevice, stream()));
        LQ_HIP_CHECK(hipMemcpyAsync(g_terms.p, g_hterms.p, j.terms_bytes, hipMemcpyHostToDevice, stream()));
        const int* d_first = (const int*)((char*)g_items.p + j.roff);
        const int* d_count = d_first + j.ninst;
        if (j.blocks > 0) {
            const uint64_t* d_x = (const uint64_t*)((const char*)g_terms.p + j.xoff);
            if (j.dtype == LQ_C128)
                hipLaunchKernelGGL(k_expect_orbits<double2>, dim3(j.blocks), dim3(kOrbitThreads), 0, stream(),
                                   (const double2*)psi, (const OrbitItem*)g_items.p, j.nitems, d_x,
                                   (const OrbitTerm*)g_terms.p, (double*)g_part.p);
            else
                hipLaunchKernelGGL(k_expect_orbits<float2>, dim3(j.blocks), dim3(kOrbitThreads), 0, stream(),
                                   (const float2*)psi, (const OrbitItem*)g_items.p, j.nitems, d_x,
                                   (const OrbitTerm*)g_terms.p, (double*)g_part.p);
            LQ_LAUNCH_CHECK();
        }
        double* d_tpart = (double*)g_part.p + j.tpart_off;
        auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
        for (size_t l = 0; l < j.lays.size(); ++l) {
            const ExpLayout& lay = j.lays[l];
            const char* dp = (const char*)g_items.p + j.lay_off[l];
            int rc = launch_expect_tile(psi, j.dtype, lay.geo, (uint64_t)j.ninst, j.dim, (const ExpTileGroup*)dp,
                                        (int)lay.groups.size(),
                                        (const ExpTileTerm*)(dp + al(lay.groups.size() * sizeof(ExpTileGroup))),
                                        (unsigned)j.tile_gx, d_tpart + l * (size_t)j.ninst * j.tile_gx);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_sum_ranges, dim3(j.ninst), dim3(256), 0, stream(), (const double*)g_part.p, d_first,
                           d_count, (const double*)d_tpart, (int)j.lays.size(), j.tile_gx, j.ninst,
                           (double*)g_out.p);
        LQ_LAUNCH_CHECK();
        LQ_HIP_CHECK(hipMemcpyAsync(g_hout.p, g_out.p, j.ninst * sizeof(double), hipMemcpyDeviceToHost, stream()));
    }
    return LQ_OK;
}

// hipGraphs of job launch sequences, keyed by everything the launches take
// as arguments (sizes, pass geometry, buffer addresses).  A key seen for the
// first time runs eagerly; the second time it is captured; later it replays.
static bool g_graphs_on = true;
static long g_graph_captures = 0, g_graph_replays = 0;
struct GraphEntry {
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
    long last_use = 0;
};
static std::vector<GraphEntry> g_graphs;
static std::vector<std::vector<uint64_t>> g_seen_keys;
static long g_graph_clock = 0;

static std::vector<uint64_t> job_key(const BatchJob& j, const void* psi) {
    std::vector<uint64_t> k = {(uint64_t)j.n, (uint64_t)j.dtype, (uint64_t)j.ninst, j.nops, j.goff, j.plan_bytes,
                               (uint64_t)j.nitems, (uint64_t)j.blocks, j.roff, j.items_bytes, j.terms_bytes, j.xoff,
                               (uint64_t)j.lays.size(), (uint64_t)j.tile_gx, j.tpart_off,
                               (uint64_t)psi, (uint64_t)g_bplan.p, (uint64_t)g_hplan.p, (uint64_t)g_items.p,
                               (uint64_t)g_hitems.p, (uint64_t)g_terms.p, (uint64_t)g_hterms.p, (uint64_t)g_part.p,
                               (uint64_t)g_out.p, (uint64_t)g_hout.p, j.plan->steps.size()};
    for (const Step& s : j.plan->steps) {
        const TilePass& p = s.pass;
        k.push_back(s.normalize ? 1 : 0);
        k.push_back(((uint64_t)(uint32_t)p.L << 32) | (uint32_t)p.hb);
        for (int h = 0; h < p.hb; ++h) k.push_back((uint64_t)p.hpos[h]);
        k.push_back(((uint64_t)(uint32_t)p.op_first << 32) | (uint32_t)p.op_count);
        k.push_back(((uint64_t)(uint32_t)p.group_first << 32) | (uint32_t)p.group_count);
        k.push_back(((uint64_t)(uint32_t)p.rich << 32));
        k.push_back(p.ntiles);
        k.push_back((uint64_t)(uintptr_t)s.jit);
    }
    k.push_back(j.alive.size());
    for (int a : j.alive) k.push_back((uint64_t)a);
    k.push_back(j.replicate ? 1 : 0);
    k.push_back(j.nover);
    k.push_back(j.hbytes);
    k.push_back(j.stage);
    for (const ExpLayout& l : j.lays) {
        k.push_back(((uint64_t)(uint32_t)l.geo.L << 32) | (uint32_t)l.geo.hb);
        for (int h = 0; h < l.geo.hb; ++h) k.push_back((uint64_t)l.geo.hpos[h]);
        k.push_back(((uint64_t)l.groups.size() << 32) | l.terms.size());
    }
    return k;
}

// Run a prepared job and wait for its energies (in g_hout).
static int job_run(const BatchJob& j, void* psi, double* energies) {
    bool graphable = g_graphs_on;
    for (const Step& s : j.plan->steps)
        if (s.normalize) graphable = false;
    int rc = LQ_OK;
    if (graphable) {
        std::vector<uint64_t> key = job_key(j, psi);
        GraphEntry* hit = nullptr;
        for (GraphEntry& e : g_graphs)
            if (e.key == key) hit = &e;
        if (!hit && std::find(g_seen_keys.begin(), g_seen_keys.end(), key) != g_seen_keys.end()) {
            // second sighting: capture the launch sequence
            LQ_HIP_CHECK(hipStreamSynchronize(stream()));
            LQ_HIP_CHECK(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
            rc = job_launch(j, psi);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(stream(), &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (e != hipSuccess) {
                set_error("graph capture failed: %s", hipGetErrorString(e));
                return LQ_ERR_DEVICE;
            }
            GraphEntry ne;
            ne.key = key;
            e = hipGraphInstantiate(&ne.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                set_error("graph instantiation failed: %s", hipGetErrorString(e));
                return LQ_ERR_DEVICE;
            }
            if (g_graphs.size() >= 16) {
                auto old = std::min_element(g_graphs.begin(), g_graphs.end(), [](const GraphEntry& a,
                                                                                 const GraphEntry& b) {
                    return a.last_use < b.last_use;
                });
                (void)hipGraphExecDestroy(old->exec);
                g_graphs.erase(old);
            }
            g_graphs.push_back(ne);
            hit = &g_graphs.back();
            ++g_graph_captures;
        }
        if (hit) {
            hit->last_use = ++g_graph_clock;
            LQ_HIP_CHECK(hipGraphLaunch(hit->exec, stream()));
            ++g_graph_replays;
        } else {
            g_seen_keys.push_back(key);
            if (g_seen_keys.size() > 64) g_seen_keys.erase(g_seen_keys.begin());
            rc = job_launch(j, psi);
        }
    } else {
        rc = job_launch(j, psi);
    }
    if (rc) return rc;
    LQ_HIP_CHECK(hipStreamSynchronize(stream()));
    if (j.blocks > 0 || !j.lays.empty()) std::memcpy(energies, g_hout.p, j.ninst * sizeof(double));
    else
        for (int i = 0; i < j.ninst; ++i) energies[i] = 0.0;
    return LQ_OK;
}

// Per-instance ops: the shared plan's ops with each instance's matrices.
// prims_i must have the same structure as the prims the plan was built from.
static void instance_ops(const Plan& plan, const std::vector<Prim>& prims_i, const std::vector<int>& first_prim,
                         TileOp* out) {
    for (size_t o = 0; o < plan.ops.size(); ++o) {
        TileOp t = plan.ops[o];
        const Prim& p = prims_i[first_prim[plan.op_src[o]]];
        std::memcpy(t.m, p.m, sizeof(t.m));
        out[o] = t;
    }
}

static std::vector<int> first_prims(const std::vector<Prim>& prims, int num_gates) {
    std::vector<int> f(num_gates, -1);
    for (size_t k = 0; k < prims.size(); ++k)
        if (prims[k].kind >= 0 && f[prims[k].src] < 0) f[prims[k].src] = (int)k;
    return f;
}

// Instances per chunk so that chunk states fit a budget of device memory.
static int chunk_instances(uint64_t dim, size_t amp, int total) {
    size_t freeb = 0, totb = 0;
    if (hipMemGetInfo(&freeb, &totb) != hipSuccess) freeb = (size_t)8 << 30;
    size_t budget = std::min<size_t>(freeb / 2, (size_t)96 << 30);
    size_t per = (size_t)dim * amp;
    int k = (int)std::max<size_t>(1, budget / std::max<size_t>(per, 1));
    return std::min(k, total);
}

int param_shift_batched(int n, int dtype, const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                        double shift, const PauliSum& ps, double* energy_out, double* grad_out) {
    const uint64_t dim = 1ull << n;
    const size_t amp = dtype == LQ_C128 ? 16 : 8;
    const int ninst = 2 * num_params + 1;
    // structure from the base circuit
    std::vector<lq_gate> bound;
    bind_template(tmpl, num_gates, params, num_params, bound);
    std::vector<Prim> base;
    int rc = gates_to_prims(n, bound.data(), num_gates, base);
    if (rc) return rc;
    Plan plan;
    plan_passes(n, base, default_tile_bits(n, dtype), default_low_bits(n, dtype), plan);
    jit_prepare(plan, dtype, (uint64_t)ninst);
    std::vector<int> fp = first_prims(base, num_gates);
    // which gates use each parameter slot
    std::vector<std::vector<int>> users(num_params);
    for (int g = 0; g < num_gates; ++g)
        for (int k = 0; k < 3; ++k)
            if (tmpl[g].param[k] >= 0 && tmpl[g].param[k] < num_params) {
                auto& u = users[tmpl[g].param[k]];
                if (u.empty() || u.back() != g) u.push_back(g);
            }
    const size_t nops = plan.ops.size();
    std::vector<int> op_of(num_gates, -1);  // each gate is one op of the plan
    for (size_t oo = 0; oo < nops; ++oo) op_of[plan.op_src[oo]] = (int)oo;
    std::vector<double> energies(ninst);
    const int chunk = chunk_instances(dim, amp, ninst);
    std::vector<PauliSum> sums(1, ps);
    std::vector<double> p(params, params + num_params);
    std::vector<Prim> one;
    // Shared prefixes (one chunk only): the circuit shifted in parameter i
    // equals instance 0's up to the step holding the first op of a gate
    // that uses it.  Instances run in the order of that step, each joins as
    // a copy of instance 0's state there, and every step runs over the
    // instances that have left instance 0's circuit by then.
    const size_t nsteps = plan.steps.size();
    std::vector<int> perm(ninst);
    for (int i = 0; i < ninst; ++i) perm[i] = i;
    std::vector<int> alive;
    if (chunk >= ninst && nsteps > 0) {
        std::vector<int> step_of_op(nops, (int)nsteps);
        for (size_t k = 0; k < nsteps; ++k)
            if (!plan.steps[k].normalize)
                for (int o = 0; o < plan.steps[k].pass.op_count; ++o)
                    step_of_op[plan.steps[k].pass.op_first + o] = (int)k;
        std::vector<int> dstep(ninst, -1);  // instance 0 leads
        for (int i = 0; i < num_params; ++i) {
            int d = (int)nsteps;
            for (int g : users[i])
                if (op_of[g] >= 0) d = std::min(d, step_of_op[op_of[g]]);
            dstep[2 * i + 1] = dstep[2 * i + 2] = d;
        }
        std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return dstep[a] < dstep[b]; });
        alive.assign(nsteps + 1, 1);
        for (int i = 1; i < ninst; ++i)
            for (size_t k = (size_t)dstep[i]; k <= nsteps; ++k) alive[k]++;
    }
    for (int c0 = 0; c0 < ninst; c0 += chunk) {
        const int cn = std::min(chunk, ninst - c0);
        BatchJob job;
        job.n = n;
        job.dtype = dtype;
        job.ninst = cn;
        job.dim = dim;
        job.plan = &plan;
        job.nops = nops;
        if (cn == ninst) job.alive = alive;
        auto shifted_prims = [&](int inst, int g, std::vector<Prim>& out) {
            // instance 2i+1: theta_i + s, instance 2i+2: theta_i - s
            const int i = (inst - 1) / 2;
            const double sh = (inst % 2 == 1) ? shift : -shift;
            p[i] = params[i] + sh;
            lq_gate gg = tmpl[g];
            for (int kk = 0; kk < 3; ++kk)
                if (gg.param[kk] >= 0) gg.theta[kk] = p[gg.param[kk]];
            p[i] = params[i];
            out.clear();
            return gates_to_prims(n, &gg, 1, out);
        };
        if (cn == ninst) {
            // one batch: instance 0's table and each instance's shifted ops
            size_t nover = 0;
            for (int k = 1; k < cn; ++k)
                for (int g : users[(perm[k] - 1) / 2]) nover += op_of[g] >= 0;
            TileOp* base_ops = nullptr;
            OpOverride* over = nullptr;
            rc = job_ops_compact(job, nover, &base_ops, &over);
            if (rc) return rc;
            std::memcpy(base_ops, plan.ops.data(), nops * sizeof(TileOp));
            size_t q = 0;
            for (int k = 1; k < cn; ++k)
                for (int g : users[(perm[k] - 1) / 2]) {
                    if (op_of[g] < 0) continue;
                    rc = shifted_prims(perm[k], g, one);
                    if (rc) return rc;
                    over[q].inst = k;
                    over[q].op = op_of[g];
                    std::memcpy(over[q].m, one[0].m, sizeof(over[q].m));
                    ++q;
                }
            std::vector<int> obs_of(cn, 0);
            std::vector<double> e(cn);
            rc = job_expect(job, sums, obs_of);
            if (!rc) rc = g_states.get((size_t)cn * dim * amp);
            if (!rc) rc = job_run(job, g_states.p, e.data());
            if (rc) return rc;
            for (int k = 0; k < cn; ++k) energies[perm[c0 + k]] = e[k];
            continue;
        }
        TileOp* ops = nullptr;
        rc = job_ops(job, &ops);
        if (rc) return rc;
        for (int k = 0; k < cn; ++k) {
            const int inst = perm[c0 + k];
            TileOp* o = ops + nops * (size_t)k;
            std::memcpy(o, plan.ops.data(), nops * sizeof(TileOp));
            if (inst == 0) continue;
            // instance 2i+1: theta_i + s, instance 2i+2: theta_i - s
            const int i = (inst - 1) / 2;
            const double sh = (inst % 2 == 1) ? shift : -shift;
            p[i] = params[i] + sh;
            for (int g : users[i]) {
                lq_gate gg = tmpl[g];
                for (int kk = 0; kk < 3; ++kk)
                    if (gg.param[kk] >= 0) gg.theta[kk] = p[gg.param[kk]];
                one.clear();
                rc = gates_to_prims(n, &gg, 1, one);
                if (rc) return rc;
                if (op_of[g] >= 0) std::memcpy(o[op_of[g]].m, one[0].m, sizeof(o[op_of[g]].m));
            }
            p[i] = params[i];
        }
        std::vector<int> obs_of(cn, 0);
        std::vector<double> e(cn);
        rc = job_expect(job, sums, obs_of);
        if (!rc) rc = g_states.get((size_t)cn * dim * amp);
        if (!rc) rc = job_run(job, g_states.p, e.data());
        if (rc) return rc;
        for (int k = 0; k < cn; ++k) energies[perm[c0 + k]] = e[k];
    }
    (void)fp;
    if (energy_out) *energy_out = energies[0];
    for (int i = 0; i < num_params; ++i) grad_out[i] = (energies[2 * i + 1] - energies[2 * i + 2]) / 2.0;
    return LQ_OK;
}

// Structure key of a circuit: kinds, qubits, flags (angles excluded).
static std::vector<int64_t> structure_key(int n, const lq_gate* g, int count) {
    std::vector<int64_t> k;
    k.reserve(2 + (size_t)count * 5);
    k.push_back(n);
    k.push_back(count);
    for (int i = 0; i < count; ++i) {
        k.push_back(g[i].kind);
        k.push_back(g[i].qubits[0]);
        k.push_back(g[i].qubits[1]);
        k.push_back(g[i].qubits[2]);
        k.push_back(g[i].flags);
        if (g[i].kind == LQ_GATE_MATRIX1 || g[i].kind == LQ_GATE_MATRIX2) k.push_back(-1000 - i);  // no sharing
    }
    return k;
}

}  // namespace lq

using namespace lq;

extern "C" int lq_batch_expectations(int count, const int* num_qubits, const lq_gate* const* circuits,
                                     const int* num_gates, const uint8_t* const* paulis,
                                     const double* const* coeffs, const int* num_terms, int convention, int dtype,
                                     double* energies) {
    LQ_API_GUARD;
    clear_error();
    if (count < 0 || (count > 0 && (!num_qubits || !circuits || !num_gates || !paulis || !coeffs || !num_terms ||
                                    !energies))) {
        set_error("invalid batch arguments");
        return LQ_ERR_INVALID_ARGUMENT;
    }
    if (dtype != LQ_C128 && dtype != LQ_C64) {
        set_error("unknown dtype %d", dtype);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    int rc = ensure_init();
    if (rc) return rc;
    // validate and compile observables
    std::vector<PauliSum> sums(count);
    for (int c = 0; c < count; ++c) {
        if (num_qubits[c] < 1 || num_qubits[c] > 30 || num_gates[c] < 0 || (num_gates[c] > 0 && !circuits[c])) {
            set_error("circuit %d: invalid size or gate list", c);
            return LQ_ERR_INVALID_ARGUMENT;
        }
        for (int g = 0; g < num_gates[c]; ++g) {
            rc = validate_gate(num_qubits[c], &circuits[c][g]);
            if (rc) return rc;
        }
        rc = compile_pauli_sum(num_qubits[c], paulis[c], coeffs[c], num_terms[c], convention, &sums[c]);
        if (rc) return rc;
    }
    // group circuits of one structure; largest states first
    std::map<std::vector<int64_t>, std::vector<int>> groups;
    for (int c = 0; c < count; ++c) groups[structure_key(num_qubits[c], circuits[c], num_gates[c])].push_back(c);
    std::vector<std::vector<int>> order;
    for (auto& kv : groups) order.push_back(kv.second);
    std::sort(order.begin(), order.end(),
              [&](const std::vector<int>& a, const std::vector<int>& b) { return num_qubits[a[0]] > num_qubits[b[0]]; });
    const size_t amp = dtype == LQ_C128 ? 16 : 8;
    for (const auto& members : order) {
        const int n = num_qubits[members[0]];
        const int ng = num_gates[members[0]];
        const uint64_t dim = 1ull << n;
        std::vector<Prim> base;
        rc = gates_to_prims(n, circuits[members[0]], ng, base);
        if (rc) return rc;
        Plan plan;
        plan_passes(n, base, default_tile_bits(n, dtype), default_low_bits(n, dtype), plan);
        jit_prepare(plan, dtype, (uint64_t)members.size());
        std::vector<int> fp = first_prims(base, ng);
        const size_t nops = plan.ops.size();
        const int total = (int)members.size();
        const int chunk = chunk_instances(dim, amp, total);
        for (int c0 = 0; c0 < total; c0 += chunk) {
            const int cn = std::min(chunk, total - c0);
            BatchJob job;
            job.n = n;
            job.dtype = dtype;
            job.ninst = cn;
            job.dim = dim;
            job.plan = &plan;
            job.nops = nops;
            TileOp* ops = nullptr;
            rc = job_ops(job, &ops);
            if (rc) return rc;
            std::vector<PauliSum> csums;
            std::vector<int> obs_of(cn);
            for (int k = 0; k < cn; ++k) {
                const int c = members[c0 + k];
                std::vector<Prim> pr;
                rc = gates_to_prims(n, circuits[c], ng, pr);
                if (rc) return rc;
                instance_ops(plan, pr, fp, ops + nops * (size_t)k);
                csums.push_back(sums[c]);
                obs_of[k] = k;
            }
            std::vector<double> e(cn);
            rc = job_expect(job, csums, obs_of);
            if (!rc) rc = g_states.get((size_t)cn * dim * amp);
            if (!rc) rc = job_run(job, g_states.p, e.data());
            if (rc) return rc;
            for (int k = 0; k < cn; ++k) energies[members[c0 + k]] = e[k];
        }
    }
    return LQ_OK;
}

extern "C" int lq_set_batch_graphs(int enable) {
    LQ_API_GUARD;
    clear_error();
    g_graphs_on = enable != 0;
    return LQ_OK;
}

extern "C" int lq_batch_graph_stats(long long* captures, long long* replays) {
    LQ_API_GUARD;
    clear_error();
    if (captures) *captures = g_graph_captures;
    if (replays) *replays = g_graph_replays;
    return LQ_OK;
}
