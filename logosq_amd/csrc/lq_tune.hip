// Autotuner: picks the per-gate kernel variant (work units per wave,
// nontemporal access) and the fused tile geometry (tile bits, contiguous low
// bits) by timing candidates with HIP events on the library stream.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lq_fused.h"

namespace lq {

int set_fused_config_internal(int tile_c128, int tile_c64, int low);
void get_fused_config(int* tile_c128, int* tile_c64, int* low);

static int time_ms(hipEvent_t a, hipEvent_t b, float* ms) {
    LQ_HIP_CHECK(hipEventSynchronize(b));
    LQ_HIP_CHECK(hipEventElapsedTime(ms, a, b));
    return LQ_OK;
}

static lq_gate mk_gate(int kind, int q0, int q1, double th) {
    lq_gate g;
    std::memset(&g, 0, sizeof(g));
    g.kind = kind;
    g.qubits[0] = q0;
    g.qubits[1] = q1;
    g.qubits[2] = -1;
    g.theta[0] = th;
    g.param[0] = g.param[1] = g.param[2] = -1;
    return g;
}

}  // namespace lq

using namespace lq;

extern "C" int lq_set_gate_kernel_config(int unroll, int nontemporal) {
    LQ_API_GUARD;
    clear_error();
    return set_gate_kernel_config(unroll, nontemporal);
}

extern "C" int lq_set_gate_unit_order(int classes) {
    LQ_API_GUARD;
    clear_error();
    return set_gate_unit_order(classes);
}

extern "C" int lq_autotune(int num_qubits, int dtype, char* report, size_t report_len) {
    LQ_API_GUARD;
    clear_error();
    if (num_qubits < 8 || num_qubits > kMaxQubits || (dtype != LQ_C128 && dtype != LQ_C64)) {
        set_error("autotune needs 8 <= num_qubits <= %d and a valid dtype", kMaxQubits);
        return LQ_ERR_INVALID_ARGUMENT;
    }
    lq_state* st = nullptr;
    int rc = lq_state_create(num_qubits, dtype, &st);
    if (rc) return rc;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        lq_state_destroy(st);
        set_error("hipEventCreate failed");
        return LQ_ERR_DEVICE;
    }
    const int n = num_qubits;
    std::string rep = "{\"num_qubits\": " + std::to_string(n) + ", \"dtype\": \"" +
                      (dtype == LQ_C128 ? "c128" : "c64") + "\", \"gate_variants\": [";
    // ---- per-gate kernels: RX on a low, a middle and the top bit; two CNOTs
    std::vector<lq_gate> probe = {mk_gate(LQ_GATE_RX, n - 1, -1, 0.3), mk_gate(LQ_GATE_RX, n / 2, -1, 0.7),
                                  mk_gate(LQ_GATE_RX, 0, -1, 1.1), mk_gate(LQ_GATE_CNOT, n / 2, n / 2 + 1, 0),
                                  mk_gate(LQ_GATE_CNOT, 1, 2, 0), mk_gate(LQ_GATE_RZ, 3, -1, 0.2)};
    const int unrolls[4] = {1, 2, 4, 8};
    float best = 1e30f;
    int best_u = 4, best_nt = 0;
    bool first = true;
    for (int nt = 0; nt < 2 && !rc; ++nt)
        for (int ui = 0; ui < 4 && !rc; ++ui) {
            set_gate_kernel_config(unrolls[ui], nt);
            float ms = 0.f;
            for (int rep_i = 0; rep_i < 2 && !rc; ++rep_i) {  // first round warms up
                (void)hipEventRecord(e0, stream());
                for (auto& g : probe) rc = rc ? rc : apply_gate_desc(st, &g);
                (void)hipEventRecord(e1, stream());
                if (!rc) rc = time_ms(e0, e1, &ms);
            }
            if (rc) break;
            char buf[128];
            snprintf(buf, sizeof(buf), "%s{\"unroll\": %d, \"nontemporal\": %d, \"ms\": %.4f}", first ? "" : ", ",
                     unrolls[ui], nt, ms);
            rep += buf;
            first = false;
            if (ms < best) {
                best = ms;
                best_u = unrolls[ui];
                best_nt = nt;
            }
        }
    if (!rc) set_gate_kernel_config(best_u, best_nt);
    {
        char buf[128];
        snprintf(buf, sizeof(buf), "], \"gate_choice\": {\"unroll\": %d, \"nontemporal\": %d}, \"unit_orders\": [",
                 best_u, best_nt);
        rep += buf;
    }
    // ---- work-unit order of the per-gate kernels: plain or transposed for all classes
    {
        const int orders[2] = {0, LQ_UNITS_ALL};
        float besto = 1e30f;
        int best_order = LQ_UNITS_ALL;
        for (int oi = 0; oi < 2 && !rc; ++oi) {
            set_gate_unit_order(orders[oi]);
            float ms = 0.f;
            for (int rep_i = 0; rep_i < 2 && !rc; ++rep_i) {
                (void)hipEventRecord(e0, stream());
                for (auto& g : probe) rc = rc ? rc : apply_gate_desc(st, &g);
                (void)hipEventRecord(e1, stream());
                if (!rc) rc = time_ms(e0, e1, &ms);
            }
            if (rc) break;
            char buf[96];
            snprintf(buf, sizeof(buf), "%s{\"classes\": %d, \"ms\": %.4f}", oi ? ", " : "", orders[oi], ms);
            rep += buf;
            if (ms < besto) {
                besto = ms;
                best_order = orders[oi];
            }
        }
        if (!rc) set_gate_unit_order(best_order);
        char buf[128];
        snprintf(buf, sizeof(buf), "], \"unit_order_choice\": %d, \"tile_variants\": [", best_order);
        rep += buf;
    }
    // ---- fused tiles: a depth-20 brickwork of RX/RY/RZ + CNOT layers (the
    // planner's windows run several layers per pass, so a shallow sample
    // ranks the geometries differently from deep circuits: depth 4 and 12
    // chose 2^13 / 4 bits, 137 ms on the 30-qubit depth-20 brickwork against
    // 131 ms at 2^12 / 4)
    std::vector<lq_gate> circ;
    for (int layer = 0; layer < 20; ++layer) {
        for (int q = 0; q < n; ++q)
            circ.push_back(mk_gate(LQ_GATE_RX + ((q + layer) % 3), q, -1, 0.1 * (q + 1) + layer));
        for (int q = layer % 2; q + 1 < n; q += 2) circ.push_back(mk_gate(LQ_GATE_CNOT, q, q + 1, 0));
    }
    int t0c, t0s, l0;
    get_fused_config(&t0c, &t0s, &l0);
    // (2^10 tiles and 6 low bits never won at 20-30 qubits: not timed)
    const int tiles128[3] = {11, 12, 13};
    const int tiles64[3] = {11, 12, 13};
    const int lows[2] = {4, 5};
    float bestf = 1e30f;
    int bt = dtype == LQ_C128 ? t0c : t0s, bl = l0;
    first = true;
    for (int ti = 0; ti < 3 && !rc; ++ti) {
        for (int li = 0; li < 2 && !rc; ++li) {
            int tb = dtype == LQ_C128 ? tiles128[ti] : tiles64[ti];
            if (tb > n) continue;
            if (dtype == LQ_C128) rc = set_fused_config_internal(tb, t0s, lows[li]);
            else rc = set_fused_config_internal(t0c, tb, lows[li]);
            if (rc) break;
            // the kernels this geometry runs are compiled before timing
            rc = jit_warm(n, dtype, circ.data(), (int)circ.size());
            if (rc) break;
            float ms = 0.f;
            for (int rep_i = 0; rep_i < 2 && !rc; ++rep_i) {
                (void)hipEventRecord(e0, stream());
                rc = execute_fused(st, circ.data(), (int)circ.size());
                (void)hipEventRecord(e1, stream());
                if (!rc) rc = time_ms(e0, e1, &ms);
            }
            if (rc) break;
            // passes of this geometry's plan and their rate (each reads and writes the state once)
            std::vector<Prim> prims;
            int passes = 0;
            if (gates_to_prims(n, circ.data(), (int)circ.size(), prims) == LQ_OK) {
                Plan plan;
                plan_passes(n, prims, tb, lows[li], plan);
                passes = plan.passes;
            }
            const double gbs = ms > 0.f ? passes * 2.0 * (dtype == LQ_C128 ? 16.0 : 8.0) * (double)(1ull << n) / ms / 1e6
                                        : 0.0;
            char buf[192];
            snprintf(buf, sizeof(buf),
                     "%s{\"tile_bits\": %d, \"low_bits\": %d, \"ms\": %.4f, \"passes\": %d, \"pass_GBps\": %.1f}",
                     first ? "" : ", ", tb, lows[li], ms, passes, gbs);
            rep += buf;
            first = false;
            if (ms < bestf) {
                bestf = ms;
                bt = tb;
                bl = lows[li];
            }
        }
    }
    if (!rc) {
        if (dtype == LQ_C128) rc = set_fused_config_internal(bt, t0s, bl);
        else rc = set_fused_config_internal(t0c, bt, bl);
    } else {
        set_fused_config_internal(t0c, t0s, l0);
    }
    {
        char buf[128];
        snprintf(buf, sizeof(buf), "], \"tile_choice\": {\"tile_bits\": %d, \"low_bits\": %d}}", bt, bl);
        rep += buf;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    lq_state_destroy(st);
    if (report && report_len) {
        std::strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return rc;
}
