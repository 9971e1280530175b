// Host-side state of the bracket-position network, shared by the MLP and
// batch translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "mlp_layout.hpp"
#include "rfd.h"
#include "rfd_internal.hpp"

struct rfd_mlp {
    rfd_ctx *ctx = nullptr;
    float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t *wb = nullptr;        // BF16 W1..W4
    int cap = 0;                   // rows the activation buffers hold
    uint16_t *x0 = nullptr;        // [cap x 400] BF16 input
    uint16_t *act[rfd::MLP_LAYERS] = {nullptr, nullptr, nullptr, nullptr};  // [cap x 256]
    uint16_t *dz[2] = {nullptr, nullptr};                                    // [cap x 256]
    float *p = nullptr, *dzh = nullptr, *loss = nullptr, *labels = nullptr;  // [cap]
    long long step = 0;
    double tau = 0.0;              // certification policy: |f| within tau of an end -> one decision
    double delta = 0.0;            // middle predictions: half-width of the bracket (0: none)
};

namespace rfd_host {

// Grows the activation buffers to hold `rows` rows.
rfd_status mlp_reserve(rfd_mlp *n, int rows);
// Forward pass on a BF16 input already on the device ([rows x 400]); writes
// p_out[rows] (sigmoid outputs) on the context's stream.  Capturable.
void mlp_forward_dev(rfd_mlp *n, const uint16_t *x0, int rows, float *p_out);

}  // namespace rfd_host
