#pragma once
#include <vector>

#include "tg_internal.h"

struct tg_mlp {
    hipStream_t stream = nullptr;
    int layers = 0;              // linear layers
    std::vector<int> dims;       // layers + 1 widths
    int width_max = 0;
    std::vector<int64_t> w_off, b_off;
    int64_t count = 0;           // parameters
    tg::DevBuf params, grads, adam_m, adam_v, zeros, acts;
    int64_t capacity = 0, last_batch = 0;
};

namespace tg {
int mlp_forward(tg_mlp *m, const float *x, int64_t batch, float *y);
int mlp_backward(tg_mlp *m, const float *dy);
}  // namespace tg
