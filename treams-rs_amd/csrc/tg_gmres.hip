// Restarted GMRES on the device, after linalg/gmres.rs:119-347: per column an
// Arnoldi basis with two Gram-Schmidt passes (here classical, batched into two
// dot/update kernels per pass), complex Givens rotations on the host, and
// convergence judged only on the recomputed true residual b - A x.  All columns
// advance in lockstep so one operator launch serves every column (the stored
// coupling is streamed once per step whatever the column count); finished columns
// carry zero vectors.  Right preconditioning keeps residual norms unpreconditioned.
#include <cmath>
#include <complex>
#include <cstdio>
#include <functional>
#include <vector>

#include "tg_gmres.h"

namespace tg {

using zc = std::complex<double>;

static int check_finite_norms(const std::vector<double> &v) {
    for (double x : v)
        if (!std::isfinite(x)) return 0;
    return 1;
}

int gmres_run(tg_cluster *c, const OpFn &A, const OpFn *Minv, const cd *b, cd *x, int cols, const GmresParams &prm,
              std::vector<tg_convergence> &conv, int *total_iterations) {
    hipStream_t s = c->stream;
    const int64_t n = c->n;
    const int restart = (int)std::min<int64_t>(std::min(prm.restart, prm.max_iterations), n);
    // buffers: V [restart+1][cols][n] (basis vector j of every column contiguous, so
    // a step's operator input is V[j] itself and its normalisation one launch), W, X
    // packed [cols][n], R residual [cols][n]
    const int64_t csv = n, vstep = (int64_t)cols * n;
    TG_TRY(c->ws[2].ensure(sizeof(cd) * (size_t)(restart + 1) * vstep));
    TG_TRY(c->ws[3].ensure(sizeof(cd) * (size_t)cols * n * 3));
    // device scalars: solution coefficients, both Gram-Schmidt passes' projections and
    // |w|^2 of an iteration, then the column scales and the extending mask
    const size_t nproj = (size_t)cols * (2 * (restart + 1) + 1);
    TG_TRY(c->ws[5].ensure(sizeof(cd) * ((size_t)cols * (restart + 2) + nproj) + sizeof(double) * 2 * cols));
    cd *V = c->ws[2].as<cd>();
    cd *W = c->ws[3].as<cd>();
    cd *X = W + (size_t)cols * n;
    cd *R = X + (size_t)cols * n;
    cd *coef = c->ws[5].as<cd>();
    cd *proj = coef + (size_t)cols * (restart + 2);
    double *scale = reinterpret_cast<double *>(proj + nproj);
    double *mask = scale + cols;
    // pinned host side: the projections and |w|^2 of a step (written there by the last
    // reduction kernel), then the column scales and the extending mask (sent up with one
    // asynchronous copy)
    TG_TRY(c->pin.ensure(sizeof(cd) * nproj + sizeof(double) * 2 * cols));

    std::vector<double> bnorm(cols), tol(cols), rn(cols);
    TG_TRY(vec_norms(c, b, n, cols, n, bnorm.data()));
    conv.assign(cols, tg_convergence{0, 0, 0.0, 0.0});
    for (int k = 0; k < cols; ++k) {
        tol[k] = std::max(prm.atol, prm.rtol * bnorm[k]);
        conv[k].rhs_norm = bnorm[k];
    }
    TG_HIP(hipMemsetAsync(x, 0, sizeof(cd) * cols * n, s));
    TG_HIP(hipMemcpyAsync(R, b, sizeof(cd) * cols * n, hipMemcpyDeviceToDevice, s));
    int iterations = 0;
    std::vector<zc> H((size_t)cols * (restart + 1) * restart), g((size_t)cols * (restart + 1));
    std::vector<double> cs((size_t)cols * restart);
    std::vector<zc> sn((size_t)cols * restart);
    std::vector<int> used(cols), extending(cols), cycle(cols);
    std::vector<cd> hbuf((size_t)cols * (restart + 1));
    std::vector<double> nxt(cols);
    // Relaxed inner applies (inexact Krylov: Simoncini and Szyld, SIAM J. Sci. Comput.
    // 25 (2003) 454; Bouras and Fraysse, SIAM J. Matrix Anal. Appl. 26 (2005) 660): once
    // the residual estimate |g_j| of every extending column is small, step j's operator
    // may err by tol / |g_j| relative without moving the attainable residual above tol.
    // The bound goes to the cluster before each apply of an inner solve; the operator
    // picks its arithmetic from it, and every other apply sees 0 (exact).
    struct RelaxReset {
        tg_cluster *c;
        ~RelaxReset() { c->relax_allowed = 0.0; }
    } relax_reset{c};
    c->relax_allowed = 0.0;
    auto relax_bound = [&](int j) {
        double allowed = INFINITY;
        for (int k = 0; k < cols; ++k) {
            if (!extending[k]) continue;
            const double gj = std::abs(g[(size_t)k * (restart + 1) + j]);
            allowed = std::min(allowed, gj > 0.0 ? tol[k] / gj : INFINITY);
        }
        return std::isfinite(allowed) ? allowed : 0.0;
    };
    // the column scales and extending masks of the next step (host side: scale_h[0 ..
    // 2 cols) in the pinned buffer) to the device array `scale`, scaling src into dst
    // when apply: by value in the scaling kernel for up to ColScales::kMax columns, else
    // by an asynchronous copy first.  Without apply nothing reads the masks before the
    // next cycle writes them, so the few-column form sends nothing.
    auto send_scales = [&](const cd *src, cd *dst, bool apply) -> int {
        const double *scale_h = reinterpret_cast<const double *>(c->pin.as<cd>() + nproj);
        if (cols <= ColScales::kMax) {
            if (!apply) return TG_OK;
            ColScales v;
            for (int k = 0; k < ColScales::kMax; ++k) {
                v.s[k] = k < cols ? scale_h[k] : 0.0;
                v.m[k] = k < cols ? scale_h[cols + k] : 0.0;
            }
            return vec_scale_cols_v(s, v, scale, src, dst, n, cols);
        }
        TG_HIP(hipMemcpyAsync(scale, scale_h, sizeof(double) * 2 * cols, hipMemcpyHostToDevice, s));
        return apply ? vec_scale_cols(s, scale, src, dst, n, cols) : TG_OK;
    };
    while (true) {
        TG_TRY(vec_norms(c, R, n, cols, n, rn.data()));
        int any = 0;
        for (int k = 0; k < cols; ++k) {
            conv[k].residual_norm = rn[k];
            conv[k].iterations = iterations;
            if (std::isfinite(rn[k]) && rn[k] <= tol[k]) {
                cycle[k] = 0;
                continue;
            }
            if (!std::isfinite(rn[k]) || iterations >= prm.max_iterations) {
                set_error("GMRES did not converge after %d iterations: residual %e, tolerance %e", iterations, rn[k],
                          tol[k]);
                if (total_iterations) *total_iterations = iterations;
                return TG_NOT_CONVERGED;
            }
            cycle[k] = 1;
            any = 1;
        }
        if (!any) break;
        // the pinned buffer (vec_norms above writes its first values only)
        cd *proj_h = c->pin.as<cd>();
        double *scale_h = reinterpret_cast<double *>(proj_h + nproj);
        // V[k][0] = r_k / |r_k| for cycling columns, 0 otherwise
        for (int k = 0; k < cols; ++k) {
            scale_h[k] = cycle[k] ? 1.0 / rn[k] : 0.0;
            scale_h[cols + k] = cycle[k] ? 1.0 : 0.0;  // the extending mask
        }
        TG_TRY(send_scales(R, V, true));
        std::fill(H.begin(), H.end(), zc(0.0));
        std::fill(g.begin(), g.end(), zc(0.0));
        for (int k = 0; k < cols; ++k) {
            g[(size_t)k * (restart + 1)] = cycle[k] ? rn[k] : 0.0;
            used[k] = 0;
            extending[k] = cycle[k];
        }
        for (int j = 0; j < restart; ++j) {
            // W = A M^{-1} V[j]
            const cd *Vj = V + (int64_t)j * vstep;
            c->relax_allowed = prm.inner ? relax_bound(j) : 0.0;
            if (Minv) {
                TG_TRY((*Minv)(Vj, R, cols));
                TG_TRY(A(R, W, cols));
            } else {
                TG_TRY(A(Vj, W, cols));
            }
            // two classical Gram-Schmidt passes against V[:, 0..j], their projections
            // (zero for columns no longer extending) kept on the device; one transfer
            // brings both passes' projections and |w|^2 to the host
            const size_t nj = (size_t)cols * (j + 1);
            cd *p0 = proj, *p1 = proj + nj;
            TG_TRY(vec_dots_dev(c, V, vstep, j + 1, W, n, cols, csv, n, p0, mask));
            TG_TRY(vec_update(s, V, vstep, j + 1, p0, W, n, cols, csv, n, -1.0));
            TG_TRY(vec_dots_dev(c, V, vstep, j + 1, W, n, cols, csv, n, p1, mask));
            TG_TRY(vec_update(s, V, vstep, j + 1, p1, W, n, cols, csv, n, -1.0));
            TG_TRY(vec_dots_dev(c, W, 0, 1, W, n, cols, n, n, nullptr, nullptr, proj_h, proj, (int64_t)(2 * nj)));
            TG_HIP(hipStreamSynchronize(s));
            for (int k = 0; k < cols; ++k)
                for (int i = 0; i <= j; ++i) {
                    zc &hh = H[((size_t)k * restart + j) * (restart + 1) + i];
                    const cd a0 = proj_h[(size_t)k * (j + 1) + i], a1 = proj_h[nj + (size_t)k * (j + 1) + i];
                    hh += zc(a0.re, a0.im);
                    hh += zc(a1.re, a1.im);
                }
            for (int k = 0; k < cols; ++k) {
                // a NaN stays NaN, as in vec_norms
                const double q = proj_h[2 * nj + k].re;
                nxt[k] = q > 0.0 ? sqrt(q) : q == q ? 0.0 : q;
            }
            int more = 0;
            for (int k = 0; k < cols; ++k) {
                double &sc = scale_h[k];
                sc = 0.0;
                if (extending[k]) {
                    zc *h = &H[((size_t)k * restart + j) * (restart + 1)];
                    h[j + 1] = nxt[k];
                    for (int i = 0; i < j; ++i) {
                        double cc = cs[(size_t)k * restart + i];
                        zc ss = sn[(size_t)k * restart + i];
                        zc a = h[i], bb = h[i + 1];
                        h[i] = cc * a + ss * bb;
                        h[i + 1] = -std::conj(ss) * a + cc * bb;
                    }
                    zc a = h[j], bb = h[j + 1];
                    double len = std::hypot(std::abs(a), std::abs(bb));
                    if (!std::isfinite(len) || len == 0.0) {
                        set_error("GMRES Arnoldi breakdown before convergence");
                        return TG_NOT_CONVERGED;
                    }
                    zc phase = std::abs(a) == 0.0 ? zc(1.0) : a / std::abs(a);
                    double cc = std::abs(a) / len;
                    zc ss = phase * std::conj(bb) / len;
                    h[j] = phase * len;
                    h[j + 1] = 0.0;
                    cs[(size_t)k * restart + j] = cc;
                    sn[(size_t)k * restart + j] = ss;
                    zc *gk = &g[(size_t)k * (restart + 1)];
                    gk[j + 1] = -std::conj(ss) * gk[j];
                    gk[j] *= cc;
                    used[k] = j + 1;
                    conv[k].iterations = iterations + 1;
                    bool stop = std::abs(gk[j + 1]) <= tol[k] || nxt[k] == 0.0 ||
                                iterations + 1 >= prm.max_iterations;
                    if (!stop) {
                        sc = 1.0 / nxt[k];
                        more = 1;
                    } else {
                        extending[k] = 0;
                    }
                }
            }
            for (int k = 0; k < cols; ++k) scale_h[cols + k] = extending[k] ? 1.0 : 0.0;
            ++iterations;
            TG_TRY(send_scales(W, V + (int64_t)(j + 1) * vstep, j + 1 < restart && more));
            if (!more) break;
        }
        // back substitution and x += M^{-1} V y, over the basis vectors written in this
        // cycle only (the rest of V holds whatever the buffer held before, and a zero
        // coefficient times a NaN there is NaN)
        int nuse = 0;
        for (int k = 0; k < cols; ++k) nuse = std::max(nuse, cycle[k] ? used[k] : 0);
        for (int k = 0; k < cols; ++k) {
            zc *gk = &g[(size_t)k * (restart + 1)];
            for (int i = used[k] - 1; i >= 0; --i) {
                zc tail = 0.0;
                for (int jj = i + 1; jj < used[k]; ++jj) tail += H[((size_t)k * restart + jj) * (restart + 1) + i] * gk[jj];
                gk[i] = (gk[i] - tail) / H[((size_t)k * restart + i) * (restart + 1) + i];
            }
            for (int i = 0; i < nuse; ++i) {
                zc v = (i < used[k] && cycle[k]) ? gk[i] : zc(0.0);
                hbuf[(size_t)k * nuse + i] = cd(v.real(), v.imag());
            }
        }
        TG_HIP(hipMemsetAsync(X, 0, sizeof(cd) * cols * n, s));
        if (nuse > 0) {
            TG_HIP(hipMemcpyAsync(coef, hbuf.data(), sizeof(cd) * cols * nuse, hipMemcpyHostToDevice, s));
            TG_TRY(vec_update(s, V, vstep, nuse, coef, X, n, cols, csv, n, 1.0));
        }
        if (Minv) {
            TG_TRY((*Minv)(X, W, cols));
            TG_TRY(vec_axpby(s, cd(1.0), W, cd(1.0), x, (int64_t)cols * n));
        } else {
            TG_TRY(vec_axpby(s, cd(1.0), X, cd(1.0), x, (int64_t)cols * n));
        }
        if (prm.inner) {
            bool all = true;
            for (int k = 0; k < cols; ++k) {
                if (!cycle[k]) continue;
                const double est = std::abs(g[(size_t)k * (restart + 1) + used[k]]);
                conv[k].residual_norm = est;
                all = all && est <= tol[k];
            }
            if (all || iterations >= prm.max_iterations) break;
        }
        // true residual R = b - A x
        c->relax_allowed = 0.0;
        TG_TRY(A(x, R, cols));
        TG_TRY(vec_axpby(s, cd(1.0), b, cd(-1.0), R, (int64_t)cols * n));
    }
    if (total_iterations) *total_iterations = iterations;
    (void)check_finite_norms;
    return TG_OK;
}

}  // namespace tg
