// Host-side helpers shared by the C ABI translation units.
#pragma once

#include <cstdint>
#include <vector>

#include "lq_common.h"
#include "lq_orbit.h"

namespace lq {

int gate_arity(int kind);
int validate_gate(int n, const lq_gate* g);
int apply_gate_desc(lq_state* st, const lq_gate* g);
void two_qubit_rot_matrix(int kind, double theta, int q1, int q2, double* M);
void gate_pair_matrix(int kind, const double* th, double* m);

// Fused (LDS tile pass) execution of a gate list (lq_fused.hip).
int execute_fused(lq_state* st, const lq_gate* gates, int count);

// A Pauli sum compiled for the device: terms grouped by their X/Y flip mask.
struct PauliTermDev {
    uint64_t zy;  // bits carrying Z or Y (sign (-1)^popcount(i & zy))
    double cr;    // coefficient * i^{#Y} (reference) or (-i)^{#Y} (physical)
    double ci;
};
struct PauliGroup {
    uint64_t x;   // flip mask (bits carrying X or Y)
    int first;    // first term in the term table
    int count;
};
struct PauliSum {
    int n = 0;
    std::vector<PauliTermDev> terms;
    std::vector<PauliGroup> groups;
};

int compile_pauli_sum(int n, const uint8_t* paulis, const double* coeffs, int num_terms, int convention,
                      PauliSum* out);

// Orbit items (lq_orbit.h): flip groups `groups` (indices into ps.groups,
// non-zero masks) packed first-fit into sets of at most maxm linearly
// independent masks, each evaluated from one read of the state; with
// diag >= 0 that group's terms (x = 0) join the first set (a set with m = 0
// when there is no flip group).
struct OrbitSet {
    int m = 0;
    int piv[kOrbitMax] = {};
    uint64_t X[1 << kOrbitMax] = {};
    std::vector<OrbitTerm> terms;
};
void plan_orbits(const PauliSum& ps, const std::vector<int>& groups, int diag, int maxm, std::vector<OrbitSet>& out);
// Groups per orbit item for a dtype: 4 (16 amplitudes per thread) for both.
int orbit_max_groups(int dtype);

// Flip groups evaluated from LDS tiles: a layout is a tile geometry (the L
// lowest bits plus hb high bits) and the flip groups whose masks lie in it.
struct ExpGeom {
    int32_t L, hb;
    int32_t hpos[28];
    uint64_t ntiles;
};
struct ExpTileGroup {
    uint32_t xl;    // flip mask in tile-local bits
    int32_t hl;     // a local bit of xl (pairs have it clear on the first row)
    int32_t first;  // terms of the group in the layout's term table
    int32_t count;
};
struct ExpTileTerm {
    uint64_t zyo;   // sign bits outside the tile
    uint32_t zyl;   // sign bits inside the tile, local positions
    int32_t im;     // 1: weight of Im(conj(psi_i) psi_j), 0: of the real part
    double g;       // 2 cr or -2 ci
};
struct ExpLayout {
    ExpGeom geo;
    std::vector<ExpTileGroup> groups;
    std::vector<ExpTileTerm> terms;
};
// Cover the sum's flip groups with tile layouts; groups that fit none (or
// all of them, for states under 8 qubits or when tiles are not wanted for
// this size) are returned in rest (indices into ps.groups).
bool expect_tiles_wanted(int n, int dtype);
// The same policy for a batch of ninst such states (auto: their total size
// beyond the 256 MB Infinity Cache).
bool expect_tiles_batch_wanted(int n, int dtype, uint64_t ninst);
void plan_expect_tiles(int n, int dtype, const PauliSum& ps, std::vector<ExpLayout>& layouts, std::vector<int>& rest);
size_t expect_tile_lds(int dtype, const ExpGeom& geo);
unsigned expect_tile_blocks(const ExpGeom& geo, uint64_t ninst);
// One launch per layout: partials [ninst][gx] of the layout's groups.
int launch_expect_tile(const void* psi, int dtype, const ExpGeom& geo, uint64_t ninst, uint64_t inst_stride,
                       const ExpTileGroup* d_groups, int ngroups, const ExpTileTerm* d_terms, unsigned gx,
                       double* d_partial);
// Expectation of a compiled sum on a state; result on the host.
int expect_pauli_sum(const lq_state* st, const PauliSum& ps, double* out);

}  // namespace lq
