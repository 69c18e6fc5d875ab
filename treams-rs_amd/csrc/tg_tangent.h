// Forward-mode derivatives (pushforward) shared by the GMRES and LU drivers.
#pragma once
#include <functional>
#include <vector>

#include "tg_internal.h"

namespace tg {

// x = (I - T C)^{-1} b for `cols` split-layout columns, filling per-column convergence.
using TangentSolve = std::function<int(const cd *b, cd *x, int cols, std::vector<tg_convergence> &conv)>;

int solve_split(tg_cluster *c, const cd *b, cd *x, int cols, const tg_gmres_options &o, int adjoint,
                std::vector<tg_convergence> &conv);

// dX for tangents of k0, the radii, permittivities, positions (host arrays; NULL = 0)
// and the incident columns (device, reference layout; NULL = 0), solved by `solve`.
int pushforward_with(tg_cluster *c, const tg_c128 *incident, const tg_c128 *x, int cols, double tangent_k0,
                     const double *tangent_radii, const tg_c128 *tangent_eps, const double *tangent_positions,
                     const tg_c128 *tangent_incident, const TangentSolve &solve, tg_c128 *dx,
                     tg_convergence *conv_host);

}  // namespace tg
