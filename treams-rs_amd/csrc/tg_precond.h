#pragma once
#include "tg_internal.h"

namespace tg {

// Group-wise preconditioners of (I - T C) (no reference counterpart: the reference
// GMRES is unpreconditioned).  Per-sphere blocks of I - T C are exactly I (C has
// zero self blocks), so block-Jacobi works on spatial groups of spheres.
int ensure_preconditioner(tg_cluster *c, int kind, int storage);
int apply_preconditioner(tg_cluster *c, const cd *x, cd *y, int cols, int adjoint);
void destroy_preconditioner(tg_cluster *c);
// free the near pairs, blocks and weights of the learned preconditioner
void learned_release(tg_cluster *c);
// rebuild near pairs, features, weights and blocks at the next use; buffers kept
void learned_invalidate(tg_cluster *c);

// Spatial groups of at most `size` spheres by recursive coordinate bisection.
std::vector<std::vector<int>> spatial_groups(const std::vector<double> &positions, int N, int size);

}  // namespace tg
