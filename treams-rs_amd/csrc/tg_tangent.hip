// Forward-mode derivatives (pushforward) of the cluster solve.
//
// Differentiating (I - T C) X = T B at the converged X (cluster/iterative.rs:362-447,
// cluster/interaction.rs:464-478, cluster/spheres.rs:263-300):
//   (I - T C) dX = dT (B + C X) + T (dB + dC X),
// with dT from the Mie blocks' derivatives along the size parameter and the
// permittivity (coeffs/mie.rs:274-305), dC X from the pair tables' tangents
// (sw/plan.rs:552-595) applied by the same subset GEMMs and plan contraction as C X,
// and one solve with the forward solver (GMRES, or the LU factors).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "tg_gmres.h"
#include "tg_internal.h"
#include "tg_tangent.h"

namespace tg {

// dT[i][l] = dT/dsize (dk0 r_i + k0 dr_i) + dT/deps deps_i on each 2x2 block.
__global__ void k_mie_tangent(const cd *mie, int N, int lmax, double k0, const double *radii, double dk0,
                              const double *dr, const double *deps, cd *out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * lmax * 4) return;
    const int i = t / (lmax * 4);
    const size_t blk = (size_t)N * lmax * 4;
    const double dsize = dk0 * radii[i] + k0 * dr[i];
    const cd de(deps[2 * i], deps[2 * i + 1]);
    out[t] = dsize * mie[blk + t] + de * mie[2 * blk + t];
}

int pushforward_with(tg_cluster *c, const tg_c128 *incident, const tg_c128 *x, int cols, double tangent_k0,
                     const double *tangent_radii, const tg_c128 *tangent_eps, const double *tangent_positions,
                     const tg_c128 *tangent_local, bool physical, const tg_c128 *tangent_incident,
                     const TangentSolve &solve, tg_c128 *dx, tg_convergence *conv_host) {
    if (!c || !incident || !x || !dx || cols < 1) {
        set_error("null argument or no columns");
        return TG_INVALID_INPUT;
    }
    if (physical) TG_TRY(require_spheres(c, "the pushforward from radii and permittivities"));
    const int N = c->N;
    // cluster/iterative.rs:378-394: finite tangents of the recorded shapes
    bool ok = std::isfinite(tangent_k0);
    for (int i = 0; ok && i < N; ++i) {
        if (tangent_radii) ok = std::isfinite(tangent_radii[i]);
        if (ok && tangent_eps) ok = std::isfinite(tangent_eps[i].re) && std::isfinite(tangent_eps[i].im);
        for (int a = 0; ok && tangent_positions && a < 3; ++a) ok = std::isfinite(tangent_positions[3 * i + a]);
    }
    if (!ok) {
        set_error("invalid illumination tangents: k0, radii, permittivities and positions must be finite");
        return TG_INVALID_INPUT;
    }
    if (!tables_fit(c)) {
        set_error("the pushforward applies dC as pair-table GEMMs: one FP64 table set (%lld bytes) must fit the "
                  "table limit",
                  (long long)((int64_t)sizeof(cd) * c->N * c->N * c->plan.T));
        return TG_INVALID_INPUT;
    }
    const int64_t count = (int64_t)cols * c->n;
    const size_t nblk = physical ? (size_t)N * c->lmax * 4 : tangent_local ? (size_t)N * c->M * c->M : 0;
    DevBuf io, tan;
    // vectors a, X, dB, R (response), W (work), rhs, dX; then dT blocks, dpos[3N], dr[N], deps[2N]
    TG_TRY(io.ensure(sizeof(cd) * (count * 7 + nblk) + sizeof(double) * 6 * (size_t)N));
    cd *a = io.as<cd>(), *X = a + count, *dB = X + count, *R = dB + count, *W = R + count, *rhs = W + count,
       *dXs = rhs + count, *dT = dXs + count;
    double *dpos = reinterpret_cast<double *>(dT + nblk), *dr = dpos + 3 * N, *deps = dr + N;
    std::vector<double> host(6 * (size_t)N, 0.0);
    for (int i = 0; i < N; ++i) {
        for (int q = 0; q < 3 && tangent_positions; ++q) host[3 * i + q] = tangent_positions[3 * i + q];
        if (tangent_radii) host[3 * N + i] = tangent_radii[i];
        if (tangent_eps) {
            host[4 * N + 2 * i] = tangent_eps[i].re;
            host[4 * N + 2 * i + 1] = tangent_eps[i].im;
        }
    }
    std::vector<tg_convergence> conv;
    int st = TG_OK;
    do {
        if (hipMemcpyAsync(dpos, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice, c->stream) !=
            hipSuccess) {
            set_error("tangent upload failed");
            st = TG_DEVICE_ERROR;
            break;
        }
        if ((st = ref_to_split(c, (const cd *)incident, a, cols))) break;
        if ((st = ref_to_split(c, (const cd *)x, X, cols))) break;
        if (tangent_incident) {
            if ((st = ref_to_split(c, (const cd *)tangent_incident, dB, cols))) break;
        } else if ((st = launch_zero_f64((double *)dB, 2 * count, c->stream))) {
            break;
        }
        // response B + C X
        if ((st = coupling_matrix_free(c, X, R, cols, 0))) break;
        if ((st = vec_axpby(c->stream, cd(1.0), a, cd(1.0), R, count))) break;
        // dB + dC X: tangent pair tables, the coupling GEMMs on them
        if ((st = tan.ensure(sizeof(cd) * (size_t)c->N * c->N * c->plan.T))) break;
        if ((st = launch_pair_tangent_tables(c, dpos, tangent_k0, tan.as<cd>()))) break;
        if ((st = coupling_tables_f64(c, tan.as<cd>(), X, W, cols))) break;
        tan.release();
        if ((st = vec_axpby(c->stream, cd(1.0), dB, cd(1.0), W, count))) break;
        // rhs = T (dB + dC X) + dT (B + C X)
        if ((st = local_apply(c, W, rhs, cols, 0, nullptr))) break;
        if (physical) {
            const double *radii = c->b_pos.as<double>() + 3 * N;
            k_mie_tangent<<<(unsigned)((nblk + 255) / 256), 256, 0, c->stream>>>(c->b_mie.as<cd>(), N, c->lmax,
                                                                                 c->k0, radii, tangent_k0, dr, deps,
                                                                                 dT);
            if (hipGetLastError() != hipSuccess) {
                set_error("Mie tangent launch failed");
                st = TG_DEVICE_ERROR;
                break;
            }
            if ((st = local_apply_blocks(c, dT, R, W, cols, 0, nullptr))) break;
        } else if (tangent_local) {
            if ((st = tloc_from_ref(c, (const cd *)tangent_local, dT))) break;
            if ((st = local_apply_dense(c, dT, R, W, cols, 0, nullptr))) break;
        }
        if ((physical || tangent_local) && (st = vec_axpby(c->stream, cd(1.0), W, cd(1.0), rhs, count))) break;
        st = solve(rhs, dXs, cols, conv);
        if (st && st != TG_NOT_CONVERGED) break;
        int s2 = split_to_ref(c, dXs, (cd *)dx, cols);
        if (!st) st = s2;
    } while (0);
    if (conv_host)
        for (int k = 0; k < cols && k < (int)conv.size(); ++k) conv_host[k] = conv[k];
    if (hipStreamSynchronize(c->stream) != hipSuccess && !st) {
        set_error("device failure in the pushforward");
        st = TG_DEVICE_ERROR;
    }
    tan.release();
    io.release();
    return st;
}

}  // namespace tg

using namespace tg;

extern "C" {

int tg_cluster_pushforward(tg_cluster *c, const tg_c128 *incident, const tg_c128 *x, int cols,
                           const tg_gmres_options *o, double tangent_k0, const double *tangent_radii,
                           const tg_c128 *tangent_eps, const double *tangent_positions,
                           const tg_c128 *tangent_incident, tg_c128 *dx, tg_convergence *conv_host) {
    if (!o) {
        set_error("null options");
        return TG_INVALID_INPUT;
    }
    tg_gmres_options od = *o;
    od.atol = 0.0;  // derivative solves are relative-only (cluster/iterative.rs:366-371)
    if (!(od.rtol > 0)) od.rtol = 1e-10;
    TangentSolve solve = [c, od](const cd *b, cd *y, int k, std::vector<tg_convergence> &conv) {
        return solve_split(c, b, y, k, od, 0, conv);
    };
    return pushforward_with(c, incident, x, cols, tangent_k0, tangent_radii, tangent_eps, tangent_positions, nullptr,
                            true, tangent_incident, solve, dx, conv_host);
}

int tg_cluster_pushforward_local(tg_cluster *c, const tg_c128 *incident, const tg_c128 *x, int cols,
                                 const tg_gmres_options *o, double tangent_k0, const double *tangent_positions,
                                 const tg_c128 *tangent_local, const tg_c128 *tangent_incident, tg_c128 *dx,
                                 tg_convergence *conv_host) {
    if (!o) {
        set_error("null options");
        return TG_INVALID_INPUT;
    }
    tg_gmres_options od = *o;
    od.atol = 0.0;
    if (!(od.rtol > 0)) od.rtol = 1e-10;
    TangentSolve solve = [c, od](const cd *b, cd *y, int k, std::vector<tg_convergence> &conv) {
        return solve_split(c, b, y, k, od, 0, conv);
    };
    return pushforward_with(c, incident, x, cols, tangent_k0, nullptr, nullptr, tangent_positions, tangent_local,
                            false, tangent_incident, solve, dx, conv_host);
}

}  // extern "C"
