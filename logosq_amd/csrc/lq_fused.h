// Fused tile-pass engine: planner output and device op format.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lq_host.h"
#include "lq_tile_types.h"

namespace lq {

// A primitive: one gate in engine terms (global bit positions).
struct Prim {
    int32_t kind;       // TileOpKind, or -1 for a renormalisation barrier
    int32_t b0, b1;     // global target bits (pair: b0; swap2/quad: b0, b1)
    int32_t sel0, sel1;
    uint64_t ctrl;
    double m[32];
    uint64_t nd;        // bits acted on non-diagonally
    uint64_t ds;        // bits the op depends on diagonally (controls, selectors)
    int32_t src;        // index of the gate this primitive came from
    int32_t mclass;     // kTPairReal for RY/CRY, kTPairRX for RX/CRX, else kTPairMat
    int32_t dclass;     // kTDiagNeg for Z/CZ, else 0
};

struct JitKernel;  // a pass kernel compiled at run time for one pass structure

struct Step {
    bool normalize;     // true: a State::normalize step, else a tile pass
    TilePass pass;
    const JitKernel* jit = nullptr;  // specialised kernel for this pass, if any
};

struct Plan {
    int n = 0;
    std::vector<Step> steps;
    std::vector<TileOp> ops;   // concatenated per pass
    std::vector<int> op_src;   // source gate of each op
    std::vector<TileGroup> groups;
    int passes = 0;
};

// Convert gates to primitives (global bit positions); angles from theta[].
int gates_to_prims(int n, const lq_gate* gates, int count, std::vector<Prim>& out);
// Plan tile passes: tiles of 2^tile_bits amplitudes whose low `low_bits`
// bits are always contiguous.
void plan_passes(int n, const std::vector<Prim>& prims, int tile_bits, int low_bits, Plan& plan);
// Tile size choice for a state (bits), given the amplitude size.
int default_tile_bits(int n, int dtype);
// The tile size for batches of states (many gates per pass).
int batch_tile_bits(int n, int dtype);
int default_low_bits(int n, int dtype);

// Launch one tile pass over `ninst` states laid out `inst_stride` amplitudes
// apart, with per-instance op tables `ops_stride` ops apart.
int launch_tile_pass(void* psi, int dtype, const TilePass& p, const TileOp* d_ops, const TileGroup* d_groups,
                     uint64_t ninst, uint64_t inst_stride, uint64_t ops_stride, const JitKernel* jit = nullptr);

// Attach specialised kernels to the plan's passes (Step::jit) as the
// run-time specialisation policy allows, compiling new pass structures;
// ninst: instances each pass will run over.  Does nothing inside a stream
// capture.
void jit_prepare(Plan& plan, int dtype, uint64_t ninst);
// Batches of circuits that start from |0...0> (the uneven batch): as
// jit_prepare for passes of any size, with the first pass's kernel (when
// steps[0].jit is set) building |0...0> in LDS instead of reading the
// state.  True when every step has a kernel.
bool jit_prepare_batch(Plan& plan, int dtype, uint64_t ninst);
// Compile (unless the mode is off) and load the specialised kernels a fused
// execution of these gates will use, so that a timed run that follows does
// not compile.
int jit_warm(int n, int dtype, const lq_gate* gates, int count);
int launch_jit(const JitKernel* k, void* psi, const TilePass& p, const TileOp* d_ops, uint64_t ninst,
               uint64_t inst_stride, uint64_t ops_stride, unsigned grid);
// The run-time specialisation policy (LQ_JIT_*), shared by the expectation
// kernels; hipRTC compile of one kernel text with the library's options
// (thread-safe); and adding to the statistics (compiles, failures, seconds).
int jit_mode();
bool jit_compile_text(const std::string& src, std::vector<char>& code);
void jit_count(int compiled, int failed, double seconds);

}  // namespace lq
