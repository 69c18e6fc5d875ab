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
// physical: T from the spheres' Mie blocks (k0, radii and permittivity tangents);
// otherwise T is an independent input with the dense tangent blocks tangent_local
// (device, N reference-layout M x M blocks; NULL = 0) and k0 enters the coupling only.
int pushforward_with(tg_cluster *c, const tg_c128 *incident, const tg_c128 *x, int cols, double tangent_k0,
                     const double *tangent_radii, const tg_c128 *tangent_eps, const double *tangent_positions,
                     const tg_c128 *tangent_local, bool physical, const tg_c128 *tangent_incident,
                     const TangentSolve &solve, tg_c128 *dx, tg_convergence *conv_host);

// Gradients with T an independent input: k0 (through the coupling), positions, the
// local blocks T-bar = Y (B + C X)^H (device, N reference-layout M x M blocks) and
// B-bar = T^H Y (device, reference layout).
int pullback_local_with(tg_cluster *c, const tg_c128 *incident, const tg_c128 *x, const tg_c128 *cotangent, int cols,
                        const TangentSolve &adjoint_solve, double *grad_k0, double *grad_positions,
                        tg_c128 *grad_local, tg_c128 *incident_grad, tg_convergence *conv_host);

}  // namespace tg
