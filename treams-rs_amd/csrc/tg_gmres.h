#pragma once
#include <functional>
#include <vector>

#include "tg_internal.h"

namespace tg {

// y = op(x) for `cols` packed split-layout columns of length n.
using OpFn = std::function<int(const cd *x, cd *y, int cols)>;

struct GmresParams {
    double rtol = 1e-10, atol = 0.0;
    int restart = 30, max_iterations = 300;
    // inner solve of a refinement loop: stop on the Arnoldi residual estimate without
    // recomputing b - A x (the caller certifies with its own FP64 residual)
    bool inner = false;
};

// y = (I - T C)^{-H} g for `cols` split-layout columns, filling per-column convergence.
using AdjointSolve = std::function<int(const cd *g, cd *y, int cols, std::vector<tg_convergence> &conv)>;

// Solve A x = b for `cols` columns from x = 0, with optional right preconditioner.
int gmres_run(tg_cluster *c, const OpFn &A, const OpFn *Minv, const cd *b, cd *x, int cols, const GmresParams &prm,
              std::vector<tg_convergence> &conv, int *total_iterations);

}  // namespace tg
