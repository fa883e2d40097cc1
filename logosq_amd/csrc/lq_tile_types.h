// Tile-pass op format shared by the host planner, the generic tile kernel
// and the kernels specialised at run time (compiled with hipRTC from this
// text, where <cstdint> is not available).
#pragma once

#ifdef __HIPCC_RTC__
using __hip_internal::int32_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#else
#include <cstdint>
#endif

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
    int32_t dclass;     // diagonal ops: kTDiagNeg when the gate kind guarantees the structure, else 0

    uint64_t ctrl;      // global bits that must be 1 (controls, fixed bits)
    uint64_t cout;      // grouped kernel: the controls outside S
    double m[32];       // pair: 8, quad: 32, diag: 8 (four phases)
};

// Specialised pair kinds used by the grouped kernel (classified at plan time).
constexpr int32_t kTPairReal = 7;  // all four entries real (RY, H-free real rotations)
constexpr int32_t kTPairRX = 8;    // real diagonal, imaginary off-diagonal (RX)
// Diagonal ops that negate the amplitudes whose controls are all set and
// leave the others (Z, CZ: m = -1 + 0i, no selectors), for every instance.
constexpr int32_t kTDiagNeg = 1;

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

}  // namespace lq
