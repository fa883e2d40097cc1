// Pauli-sum expectations with run-time specialised LDS-tile kernels
// (lq_expect_jit.hip): the plan of a sum on a state and its launch.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "lq_host.h"

namespace lq {

struct XKernel;  // one compiled layout kernel (opaque)

// A signed sum of raw coefficients: (index, +1 / -1) pairs.
typedef std::vector<std::pair<int, int>> Combo;

struct XTerm {
    int gi;         // index of its coefficient in the layout's table
    uint64_t zyo;   // sign bits outside the tile (global positions)
    uint32_t zyr;   // sign bits in the tile but outside the window (local); pair groups: all local bits
    uint32_t zw;    // sign bits in the window (window-local); pair groups: 0
    int im;         // weight of Im(conj(psi_i) psi_j) rather than of the real part
};
struct XGroup {
    uint32_t xl;  // flip mask, local bits
    uint32_t xw;  // flip mask, window-local bits (window groups)
    int hw;       // a bit of xw (window groups) or of xl (pair groups): rows with it clear
    std::vector<XTerm> terms;
};
struct XWindow {
    std::vector<int> bits;      // local positions, window bit k = bits[k]
    std::vector<int> tbits;     // local positions taken by thread index bits 0.. in order
    std::vector<XGroup> groups;
    std::vector<XTerm> diag;    // the diagonal group, on the first window of the first layout
};
struct XLayout {
    int K = 0, L = 0;
    std::vector<int> hpos;      // global positions of local bits L..K-1, ascending
    std::vector<XWindow> windows;
    std::vector<XGroup> pairs;  // groups in the tile but in no window
    std::vector<double> coef;   // per term (gi order): 2 cr, -2 ci (flip) or cr (diagonal)
    std::vector<Combo> derived; // table entries after coef: sums the kernel text uses
    XKernel* k = nullptr;
};

struct XPlan {
    int n = 0, dtype = 0;
    std::vector<XLayout> layouts;
    std::vector<int> rest;  // flip groups of the sum (indices into ps.groups) in no layout
    bool diag_in = false;   // the diagonal group is evaluated by the first layout
};

// The policy wants the specialised kernels for an n-qubit state (run-time
// specialisation not off, tiles wanted at this size, one wavefront per tile).
bool expect_jit_wanted(int n, int dtype);
// The same for a batch of ninst n-qubit states sharing one sum.
bool expect_jit_wanted_batch(int n, int dtype, uint64_t ninst);
// Plan the layouts of a compiled sum; false when the sum does not suit them.
bool plan_expect_jit(int n, int dtype, const PauliSum& ps, XPlan& xp);
// Attach compiled kernels (compiling as the policy allows; force: at first
// sight; load=false: compile only).  True when every layout has one.
bool prepare_expect_jit(XPlan& xp, bool force, bool load);
// Coefficient table of the plan (all layouts, in launch order).
size_t expect_jit_coef_count(const XPlan& xp);
void expect_jit_coefs(const XPlan& xp, double* out);
// Blocks per layout launch for ninst instances.
int expect_jit_blocks(const XPlan& xp, uint64_t ninst);
// The plan as JSON: per layout the geometry, the coefficient table and the
// kernel text (introspection for tests).
std::string describe_expect_jit(const XPlan& xp);
// One launch per layout; partials [layout][ninst][gx] into d_partial.
int launch_expect_jit(const XPlan& xp, const void* psi, const double* d_coef, uint64_t coef_stride, uint64_t ninst,
                      uint64_t inst_stride, int gx, double* d_partial);

}  // namespace lq
