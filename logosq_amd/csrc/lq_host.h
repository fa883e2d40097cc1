// Host-side helpers shared by the C ABI translation units.
#pragma once

#include <cstdint>
#include <vector>

#include "lq_common.h"

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
// Expectation of a compiled sum on a state; result on the host.
int expect_pauli_sum(const lq_state* st, const PauliSum& ps, double* out);

}  // namespace lq
