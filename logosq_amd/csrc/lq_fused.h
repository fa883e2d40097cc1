// Fused tile-pass engine: planner output and device op format.
#pragma once

#include <cstdint>
#include <vector>

#include "lq_host.h"

namespace lq {

// Device op kinds inside a tile pass.
enum TileOpKind : int32_t {
    kTPairMat = 0,   // 2x2 matrix on local bit lb0
    kTPairSwap = 1,  // X-type swap on lb0
    kTPairNegY = 2,  // the reference's YGate (-Y)
    kTPairH = 3,     // Hadamard in the reference's s*(a0 +- a1) form
    kTDiag = 4,      // phase d[bit(sel0) | bit(sel1) << 1] where ctrl bits are all 1
    kTSwap2 = 5,     // SWAP of local bits lb0 (value 1) and lb1 (value 0)
    kTQuad = 6       // 4x4 on local bits (lb0 = a, lb1 = b), sub-index (bit a << 1) | bit b
};

struct TileOp {
    int32_t kind;
    int32_t lb0, lb1;   // local (tile) bit positions
    int32_t sel0, sel1; // global bit positions for diagonal selectors (-1: none)
    int32_t osel0, osel1; // grouped kernel: selectors outside the group's subcube (-1: none / inside)
    // 32-bit fields throughout: the kernels read them with scalar loads
    int32_t sp0, sp1;   // positions of lb0 / lb1 inside the op's group subcube
    int32_t s0p, s1p;   // grouped kernel: subcube positions of selectors inside S (-1: none)
    int32_t cin;        // grouped kernel: controls inside S, as a mask of subcube positions
    int32_t mclass;     // pair ops: kTPairReal / kTPairRX when the gate kind guarantees the structure

    uint64_t ctrl;      // global bits that must be 1 (controls, fixed bits)
    uint64_t cout;      // grouped kernel: the controls outside S
    double m[32];       // pair: 8, quad: 32, diag: 8 (four phases)
};

// Specialised pair kinds used by the grouped kernel (classified at plan time).
constexpr int32_t kTPairReal = 7;  // all four entries real (RY, H-free real rotations)
constexpr int32_t kTPairRX = 8;    // real diagonal, imaginary off-diagonal (RX)

constexpr int kMaxTileHigh = 28;
constexpr int kGroupBits = 4;  // a group acts on at most 4 tile bits (16-amplitude subcubes)

// A run of consecutive ops of one pass whose non-diagonal target bits all lie
// in a set S of kGroupBits tile bits: each thread gathers one 2^4 subcube
// (all values of the S bits) into registers and applies the whole run there.
struct TileGroup {
    int32_t op_first;   // relative to the pass's first op
    int32_t op_count;
    int32_t spos[kGroupBits];  // local positions of S, ascending
    uint64_t gbit[kGroupBits]; // the corresponding global bits
};

struct TilePass {
    int32_t L;          // low contiguous bits 0..L-1 in every tile
    int32_t hb;         // number of high tile bits
    int32_t hpos[kMaxTileHigh];  // their global positions, ascending
    int32_t op_first;   // ops of this pass: [op_first, op_first + op_count)
    int32_t op_count;
    int32_t group_first; // groups of this pass (grouped kernel, K >= 8)
    int32_t group_count;
    int32_t rich;       // the pass holds SWAP / 4x4 ops (kernel variant with those paths)
    uint64_t ntiles;    // 2^(n - L - hb)
};

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
};

struct Step {
    bool normalize;     // true: a State::normalize step, else a tile pass
    TilePass pass;
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
int default_low_bits(int n, int dtype);

// Launch one tile pass over `ninst` states laid out `inst_stride` amplitudes
// apart, with per-instance op tables `ops_stride` ops apart.
int launch_tile_pass(void* psi, int dtype, const TilePass& p, const TileOp* d_ops, const TileGroup* d_groups,
                     uint64_t ninst, uint64_t inst_stride, uint64_t ops_stride);

}  // namespace lq
