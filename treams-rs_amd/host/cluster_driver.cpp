// C++ host driver over the C ABI: the call sequence a Rust host in treams-core makes
// through its FFI crate, for one plane wave on a sphere cluster.
//
//   1. dense: sphere_cluster_factor -> InteractionFactor::solve -> cross sections
//      (cluster/spheres.rs:85-94, cluster/interaction.rs:213-218);
//   2. iterative: IterativeSphereCluster::solve (cluster/iterative.rs:264-284);
//   3. gradient of sigma_sca with respect to radii and positions through the LU
//      adjoint (SphereClusterResidual::pullback, cluster/spheres.rs:321-378), the
//      regular-overlap pullback and the plane-wave phase.
//
// usage: cluster_driver LMAX K0 < spheres   (one "radius eps_re eps_im x y z" per line)
// prints one JSON object.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/treams_gpu.h"

using zc = std::complex<double>;

static int fail(const char *what, int st) {
    std::fprintf(stderr, "%s failed (status %d): %s\n", what, st, tg_last_error());
    return 1;
}

#define CHECK(call)                          \
    do {                                     \
        int _s = (call);                     \
        if (_s != TG_OK) return fail(#call, _s); \
    } while (0)

#define HIPCHECK(call)                                                           \
    do {                                                                         \
        hipError_t _e = (call);                                                  \
        if (_e != hipSuccess) {                                                  \
            std::fprintf(stderr, "%s failed: %s\n", #call, hipGetErrorString(_e)); \
            return 1;                                                            \
        }                                                                        \
    } while (0)

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s LMAX K0 < spheres (radius eps_re eps_im x y z per line)\n", argv[0]);
        return 2;
    }
    const int lmax = std::atoi(argv[1]);
    const double k0 = std::atof(argv[2]);
    std::vector<double> radii, pos;
    std::vector<tg_c128> eps;
    double r, er, ei, x, y, z;
    while (std::scanf("%lf %lf %lf %lf %lf %lf", &r, &er, &ei, &x, &y, &z) == 6) {
        radii.push_back(r);
        eps.push_back(tg_c128{er, ei});
        pos.insert(pos.end(), {x, y, z});
    }
    const int N = (int)radii.size();
    if (N == 0) {
        std::fprintf(stderr, "no spheres on stdin (radius eps_re eps_im x y z per line)\n");
        return 2;
    }
    tg_cluster *c = nullptr;
    CHECK(tg_cluster_create(lmax, k0, N, radii.data(), eps.data(), pos.data(), nullptr, &c));
    const int64_t n = tg_cluster_dimension(c);
    const int M = tg_modes_per_sphere(lmax);
    tg_c128 *a, *xl, *xg, *rx, *rhx, *g, *abar;
    for (tg_c128 **p : {&a, &xl, &xg, &rx, &rhx, &g, &abar}) HIPCHECK(hipMalloc((void **)p, sizeof(tg_c128) * n));
    const double dir[3] = {0.0, 0.0, 1.0};
    CHECK(tg_cluster_plane_wave(c, dir, 1, a));

    // 1. dense LU (c128)
    tg_factor *f = nullptr;
    tg_convergence conv;
    CHECK(tg_factor_create(c, TG_STORE_C128, &f));
    CHECK(tg_factor_solve(f, a, 1, 1e-12, 0, xl, &conv));
    double sca_lu, ext_lu;
    CHECK(tg_cluster_cross_sections(c, a, xl, 0.5, &sca_lu, &ext_lu));
    const double lu_res = conv.residual_norm / conv.rhs_norm;

    // 2. GMRES (the reference's unpreconditioned FP64 method)
    tg_gmres_options o{1e-10, 0.0, 30, 300, 0, TG_PC_NONE};
    tg_convergence cg;
    CHECK(tg_cluster_solve(c, a, 1, &o, 0, xg, &cg));
    double sca_gm, ext_gm;
    CHECK(tg_cluster_cross_sections(c, a, xg, 0.5, &sca_gm, &ext_gm));

    // 3. d sigma_sca: cotangent G = (R + R^H) x / k^2 (flux 0.5), LU adjoint pullback
    CHECK(tg_cluster_overlap(c, xl, rx, 1, 0));
    CHECK(tg_cluster_overlap(c, xl, rhx, 1, 1));
    std::vector<tg_c128> h_rx(n), h_rhx(n), h_g(n), h_x(n), h_a(n), h_ab(n), h_u(n);
    HIPCHECK(hipMemcpy(h_rx.data(), rx, sizeof(tg_c128) * n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(h_rhx.data(), rhx, sizeof(tg_c128) * n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(h_x.data(), xl, sizeof(tg_c128) * n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(h_a.data(), a, sizeof(tg_c128) * n, hipMemcpyDeviceToHost));
    const double w = 1.0 / (k0 * k0);
    for (int64_t i = 0; i < n; ++i) {
        h_g[i] = tg_c128{(h_rx[i].re + h_rhx[i].re) * w, (h_rx[i].im + h_rhx[i].im) * w};
        h_u[i] = tg_c128{h_x[i].re * w, h_x[i].im * w};
    }
    HIPCHECK(hipMemcpy(g, h_g.data(), sizeof(tg_c128) * n, hipMemcpyHostToDevice));
    double gk;
    std::vector<double> gr(N), gp(3 * N), gpr(3 * N);
    std::vector<tg_c128> ge(N);
    CHECK(tg_factor_pullback(f, a, xl, g, 1, 1e-12, &gk, gr.data(), ge.data(), gp.data(), abar, &conv));
    // regular overlap: d/dpos Re<x x^H w, R>
    HIPCHECK(hipMemcpy(rx, h_u.data(), sizeof(tg_c128) * n, hipMemcpyHostToDevice));
    CHECK(tg_cluster_overlap_pullback(c, rx, xl, 1, gpr.data()));
    // plane-wave phase e^{i k.r_j}: d a_j / d r_j = i k a_j
    HIPCHECK(hipMemcpy(h_ab.data(), abar, sizeof(tg_c128) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < N; ++i) {
        double phase = 0.0;
        for (int m = 0; m < M; ++m) {
            const zc ab(h_ab[i * M + m].re, h_ab[i * M + m].im), av(h_a[i * M + m].re, h_a[i * M + m].im);
            phase += std::real(std::conj(ab) * zc(0.0, 1.0) * av);
        }
        for (int d = 0; d < 3; ++d) gp[3 * i + d] += gpr[3 * i + d] + phase * k0 * dir[d];
    }

    std::printf("{\"n\": %lld, \"lu\": {\"scattering\": %.17g, \"extinction\": %.17g, \"relative_residual\": %.3g},\n",
                (long long)n, sca_lu, ext_lu, lu_res);
    std::printf(" \"gmres\": {\"scattering\": %.17g, \"extinction\": %.17g, \"iterations\": %d},\n", sca_gm, ext_gm,
                cg.iterations);
    std::printf(" \"grad_radii\": [");
    for (int i = 0; i < N; ++i) std::printf("%s%.17g", i ? ", " : "", gr[i]);
    std::printf("],\n \"grad_positions\": [");
    for (int i = 0; i < 3 * N; ++i) std::printf("%s%.17g", i ? ", " : "", gp[i]);
    std::printf("]}\n");

    CHECK(tg_factor_destroy(f));
    for (tg_c128 *p : {a, xl, xg, rx, rhx, g, abar}) HIPCHECK(hipFree(p));
    CHECK(tg_cluster_destroy(c));
    return 0;
}
