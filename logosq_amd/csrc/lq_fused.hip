// Fused tile-pass execution (placeholder until the tile engine lands).
#include "lq_host.h"

namespace lq {

int execute_fused(lq_state* st, const lq_gate* gates, int count) {
    (void)st;
    (void)gates;
    (void)count;
    set_error("fused execution is not built yet");
    return LQ_ERR_UNSUPPORTED;
}

int param_shift_batched(int n, int dtype, const lq_gate* tmpl, int num_gates, const double* params, int num_params,
                        double shift, const PauliSum& ps, double* energy_out, double* grad_out) {
    (void)n; (void)dtype; (void)tmpl; (void)num_gates; (void)params; (void)num_params; (void)shift; (void)ps;
    (void)energy_out; (void)grad_out;
    set_error("batched gradient is not built yet");
    return LQ_ERR_UNSUPPORTED;
}

}  // namespace lq
