// Adjoint of the cluster solve: pair-translation gradients and Mie gradients.
//
// For a real loss with dL = Re<G, dX>, X = (I - T C)^{-1} T B, the adjoint is
// Y = (I - T C)^{-H} G and (cluster/interaction.rs:533-558)
//   B-bar = T^H Y,  C-bar = T^H Y X^H,  T-bar = Y (B + C X)^H on the 2x2 blocks.
// C-bar restricted to a pair block is contracted with the conjugate plan weights
// into one cotangent per harmonic (sw/plan.rs:205-213), which pulls back through the
// derivative table of z_p(k r) Y_p^q(d/r) to the displacement and k (sw/plan.rs:596-615).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "tg_gmres.h"
#include "tg_internal.h"
#include "tg_special.h"
#include "tg_tangent.h"

namespace tg {

int solve_split(tg_cluster *c, const cd *b, cd *x, int cols, const tg_gmres_options &o, int adjoint,
                std::vector<tg_convergence> &conv);

__device__ __forceinline__ int unit_dir(const double d[3], double &r, double u[3]) {
    double s = fmax(fabs(d[0]), fmax(fabs(d[1]), fabs(d[2])));
    if (s < 2.2250738585072014e-308) return 0;
    double a = d[0] / s, b = d[1] / s, c = d[2] / s;
    double nrm = sqrt(a * a + b * b + c * c);
    r = s * nrm;
    u[0] = a / nrm; u[1] = b / nrm; u[2] = c / nrm;
    return 1;
}

// One workgroup per ordered pair (i = blockIdx.y destination, j = blockIdx.x source).
// out[(i * N + j) * 4 + {0,1,2,3}] = (g_x, g_y, g_z, Re g_k) of the pair's block
// cotangent sum_c U_i[c] V_j[c]^H (zero for i == j).
__global__ __launch_bounds__(256) void k_pair_gradient(DevPlan pl, double k, const double *pos, int N, int cols,
                                                       const cd *U, const cd *V, int64_t n, int64_t nh, int regular,
                                                       double *out, int accumulate) {
    extern __shared__ cd sm[];
    const int H = pl.H, L2 = pl.L2, T = pl.T;
    const int i = blockIdx.y, j = blockIdx.x;
    double *o = out + ((int64_t)i * N + j) * 4;
    if (i == j) {
        if (threadIdx.x < 4 && !accumulate) o[threadIdx.x] = 0.0;
        return;
    }
    cd *us = sm;                  // [cols][2][H]
    cd *vs = us + cols * 2 * H;   // [cols][2][H]
    cd *Y = vs + cols * 2 * H;    // [T]
    cd *G = Y + T;                // [T][3]
    cd *rad = G + 3 * T;          // [L2 + 2]
    __shared__ double su[3], sr;
    __shared__ double red[4][256];
    const int tid = threadIdx.x;
    for (int t = tid; t < cols * 2 * H; t += blockDim.x) {
        const int cc = t / (2 * H), rem = t - cc * 2 * H, h = rem / H, a = rem - h * H;
        us[t] = U[(int64_t)cc * n + (int64_t)h * nh + (int64_t)i * H + a];
        vs[t] = V[(int64_t)cc * n + (int64_t)h * nh + (int64_t)j * H + a];
    }
    if (tid == 0) {
        double d[3], r, u[3];
        for (int a = 0; a < 3; ++a) d[a] = pos[3 * i + a] - pos[3 * j + a];
        unit_dir(d, r, u);
        su[0] = u[0]; su[1] = u[1]; su[2] = u[2];
        sr = r;
        const double kr = k * r;
        if (!regular) {
            tg_hankel_seq(L2 + 1, kr, rad);
        } else {
            double jj[64];
            tg_bessel_j_seq(L2 + 1, kr, jj);
            for (int p = 0; p <= L2 + 1; ++p) rad[p] = cd(jj[p]);
        }
    }
    __syncthreads();
    // solid harmonics and gradients at the unit vector, one thread per order q >= 0
    for (int q = tid; q <= L2; q += blockDim.x) {
        const double x = su[0], y = su[1], z = su[2];
        cd xy(x, y);
        cd s(1.0), g0(0.0), g1(0.0), g2(0.0);
        for (int kk = 1; kk <= q; ++kk) {
            const double f = -(double)(2 * kk - 1);
            g0 = f * (g0 * xy + s);
            g1 = f * (g1 * xy + s * cd(0.0, 1.0));
            g2 = f * (g2 * xy);
            s = f * (s * xy);
        }
        cd ps(0.0), pg0(0.0), pg1(0.0), pg2(0.0);
        for (int p = q; p <= L2; ++p) {
            if (p > q) {
                const double a = (double)(2 * p - 1) / (double)(p - q);
                const double b = (double)(p + q - 1) / (double)(p - q);
                cd ns = (a * z) * s - b * ps;
                cd n0 = a * (z * g0) - b * (pg0 + (2.0 * x) * ps);
                cd n1 = a * (z * g1) - b * (pg1 + (2.0 * y) * ps);
                cd n2 = a * (z * g2 + s) - b * (pg2 + (2.0 * z) * ps);
                ps = s; pg0 = g0; pg1 = g1; pg2 = g2;
                s = ns; g0 = n0; g1 = n1; g2 = n2;
            }
            const int idx = p * p + p + q;
            Y[idx] = s;
            G[3 * idx + 0] = g0; G[3 * idx + 1] = g1; G[3 * idx + 2] = g2;
            if (q > 0) {
                const int nidx = p * p + p - q;
                const double f = pl.refl[nidx];
                Y[nidx] = f * conj(s);
                G[3 * nidx + 0] = f * conj(g0); G[3 * nidx + 1] = f * conj(g1); G[3 * nidx + 2] = f * conj(g2);
            }
        }
    }
    __syncthreads();
    double gx = 0.0, gy = 0.0, gz = 0.0, gk = 0.0;
    const double r = sr;
    for (int idx = tid; idx < T; idx += blockDim.x) {
        const int p = (int)sqrt((double)idx + 0.5);
        // contracted cotangent of harmonic idx
        cd acc(0.0);
        for (int t = pl.tstarts[idx]; t < pl.tstarts[idx + 1]; ++t) {
            const int ent = pl.tentry[t];
            const int e = ent & 0x3fffffff;
            const bool cross = (ent & 0x40000000) != 0;
            const int a = e / H, b = e - a * H;
            cd c0(0.0), c1(0.0);
            for (int cc = 0; cc < cols; ++cc) {
                c0 += us[cc * 2 * H + b] * conj(vs[cc * 2 * H + a]);
                c1 += us[cc * 2 * H + H + b] * conj(vs[cc * 2 * H + H + a]);
            }
            if (!cross) {
                acc += pl.tw[t] * (c0 + c1);
            } else {
                cd dd = c0 - c1;
                acc += pl.tw[t] * cd(dd.im, -dd.re);  // -i (c0 - c1)
            }
        }
        const double x = k * r;
        const cd zp = rad[p];
        const cd zd = ((double)p / x) * zp - rad[p + 1];  // z_p'(x)
        const cd Yv = Y[idx];
        const cd dk = (zd * r) * Yv;
        const cd scaled = zp / r;
        double u[3] = {su[0], su[1], su[2]};
        double g[3];
        for (int a = 0; a < 3; ++a) {
            cd tangent = G[3 * idx + a] - ((double)p * u[a]) * Yv;
            cd dpos = (zd * (k * u[a])) * Yv + scaled * tangent;
            g[a] = conj(acc).re * dpos.re - conj(acc).im * dpos.im;  // Re(conj(acc) dpos)
        }
        gx += g[0];
        gy += g[1];
        gz += g[2];
        gk += acc.re * dk.re + acc.im * dk.im;  // Re(acc conj(dk))
    }
    red[0][tid] = gx; red[1][tid] = gy; red[2][tid] = gz; red[3][tid] = gk;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s)
            for (int c = 0; c < 4; ++c) red[c][tid] += red[c][tid + s];
        __syncthreads();
    }
    if (tid < 4) o[tid] = (accumulate ? o[tid] : 0.0) + red[tid][0];
}

// grad_pos[i] = sum_j g(i, j) - sum_j g(j, i), in j order; gk_rows[i] = sum_j of the
// k parts of row i, in j order.
__global__ void k_pair_reduce(const double *pg, int N, double *grad_pos, double *gk_rows) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double g[3] = {0, 0, 0}, gk = 0.0;
    for (int j = 0; j < N; ++j) {
        for (int a = 0; a < 3; ++a) g[a] += pg[((int64_t)i * N + j) * 4 + a] - pg[((int64_t)j * N + i) * 4 + a];
        gk += pg[((int64_t)i * N + j) * 4 + 3];
    }
    for (int a = 0; a < 3; ++a) grad_pos[3 * i + a] = g[a];
    gk_rows[i] = gk;
}

// gk = sum_i gk_rows[i] in i order (one thread: N values).
__global__ void k_sum_in_order(const double *v, int N, double *out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < N; ++i) s += v[i];
        *out = s;
    }
}

int pair_gradients(tg_cluster *c, const cd *u, const cd *v, int cols, int regular, double *grad_pos_dev,
                   double *grad_k_dev) {
    const int H = c->H;
    DevBuf pg;
    TG_TRY(pg.ensure(sizeof(double) * (4 * (size_t)c->N * c->N + c->N)));
    double *gk_rows = pg.as<double>() + 4 * (size_t)c->N * c->N;
    // columns per launch so that the cotangent vectors and tables fit the LDS left
    // beside the kernel's static reduction arrays (8 KB + a few words); later chunks
    // accumulate into the same pair gradients (fixed order)
    const size_t fixed = sizeof(cd) * (4 * (size_t)c->plan.T + c->plan.L2 + 2);
    const size_t budget = 65536 - 8192 - 256;
    const int cmax = (int)std::max<size_t>(1, (budget - fixed) / (sizeof(cd) * 4 * H));
    Timer::Rec rec;
    TG_TRY(c->timer.begin("pair_gradient", c->stream, rec));
    for (int c0 = 0; c0 < cols; c0 += cmax) {
        const int nc = std::min(cmax, cols - c0);
        const size_t smem = fixed + sizeof(cd) * (size_t)nc * 4 * H;
        k_pair_gradient<<<dim3(c->N, c->N), 256, smem, c->stream>>>(
            c->plan, c->k0, c->b_pos.as<double>(), c->N, nc, u + (int64_t)c0 * c->n, v + (int64_t)c0 * c->n, c->n,
            c->nh, regular, pg.as<double>(), c0 > 0);
        TG_HIP(hipGetLastError());
    }
    TG_TRY(c->timer.end("pair_gradient", c->stream, rec));
    k_pair_reduce<<<(c->N + 63) / 64, 64, 0, c->stream>>>(pg.as<double>(), c->N, grad_pos_dev, gk_rows);
    k_sum_in_order<<<1, 64, 0, c->stream>>>(gk_rows, c->N, grad_k_dev);
    TG_HIP(hipGetLastError());
    TG_HIP(hipStreamSynchronize(c->stream));
    pg.release();
    return TG_OK;
}

// Mie gradients per sphere: T-bar = sum over orders and columns of Y_b (B + C X)_b^H on
// each degree's 2x2 block; g_size = Re sum conj(T-bar) dT/dx, g_eps = sum T-bar conj(dT/deps).
__global__ void k_mie_pullback(const cd *mie, int N, int lmax, int H, int64_t n, int64_t nh, int cols, const cd *Y,
                               const cd *R, double k0, const double *radii, double *grad_r, cd *grad_e,
                               double *grad_k_part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t blk = (size_t)N * lmax * 4;
    double gr = 0.0, gk = 0.0;
    cd ge(0.0);
    for (int l = 1; l <= lmax; ++l) {
        cd tb[4] = {cd(0.0), cd(0.0), cd(0.0), cd(0.0)};
        for (int m = -l; m <= l; ++m) {
            const int a = l * l + l + m - 1;
            for (int cc = 0; cc < cols; ++cc) {
                const int64_t ip = (int64_t)cc * n + (int64_t)i * H + a, im = ip + nh;
                cd y[2] = {Y[ip], Y[im]}, rr[2] = {R[ip], R[im]};
                for (int r = 0; r < 2; ++r)
                    for (int s = 0; s < 2; ++s) tb[r * 2 + s] += y[r] * conj(rr[s]);
            }
        }
        const cd *dx = mie + blk + (size_t)(i * lmax + l - 1) * 4;
        const cd *de = mie + 2 * blk + (size_t)(i * lmax + l - 1) * 4;
        double gs = 0.0;
        for (int e = 0; e < 4; ++e) {
            gs += tb[e].re * dx[e].re + tb[e].im * dx[e].im;  // Re(conj(tb) dx)
            ge += tb[e] * conj(de[e]);
        }
        gr += k0 * gs;
        gk += radii[i] * gs;
    }
    grad_r[i] = gr;
    grad_e[i] = ge;
    grad_k_part[i] = gk;
}

}  // namespace tg

using namespace tg;

namespace tg {

// Gradients through X = (I - T C)^{-1} T B given an adjoint solver (GMRES or the LU
// factors): Y = A^{-H} G, B-bar = T^H Y, the pair gradients of C-bar = T^H Y X^H and
// the Mie gradients of T-bar = Y (B + C X)^H.
int pullback_with(tg_cluster *c, const tg_c128 *incident, const tg_c128 *x, const tg_c128 *cotangent, int cols,
                  const AdjointSolve &adjoint_solve, double *grad_k0, double *grad_radii, tg_c128 *grad_eps,
                  double *grad_positions, tg_c128 *incident_grad, tg_convergence *conv_host) {
    TG_TRY(require_spheres(c, "the pullback to radii, permittivities and positions"));
    if (!c || !incident || !x || !cotangent || cols < 1) {
        set_error("null argument or no columns");
        return TG_INVALID_INPUT;
    }
    const int64_t count = (int64_t)cols * c->n;
    DevBuf io;
    // 6 vectors, then gpos[3N], gr[N], gkp[N], gk[2], ge[N] (complex)
    TG_TRY(io.ensure(sizeof(cd) * count * 6 + sizeof(double) * (5 * (size_t)c->N + 2) + sizeof(cd) * c->N));
    cd *a = io.as<cd>(), *X = a + count, *Gs = X + count, *Y = Gs + count, *U = Y + count, *Rsp = U + count;
    double *gpos = reinterpret_cast<double *>(Rsp + count);
    double *gr = gpos + 3 * c->N;
    double *gkp = gr + c->N;
    double *gk = gkp + c->N;
    cd *ge = reinterpret_cast<cd *>(gk + 2);
    int st = TG_OK;
    std::vector<tg_convergence> conv;
    do {
        if ((st = ref_to_split(c, (const cd *)incident, a, cols))) break;
        if ((st = ref_to_split(c, (const cd *)x, X, cols))) break;
        if ((st = ref_to_split(c, (const cd *)cotangent, Gs, cols))) break;
        if ((st = adjoint_solve(Gs, Y, cols, conv))) break;
        if ((st = local_apply(c, Y, U, cols, 1, nullptr))) break;  // T^H Y
        if ((st = split_to_ref(c, U, (cd *)incident_grad, cols))) break;
        if ((st = pair_gradients(c, U, X, cols, 0, gpos, gk))) break;
        // response B + C X
        if ((st = coupling_matrix_free(c, X, Rsp, cols, 0))) break;
        if ((st = vec_axpby(c->stream, cd(1.0), a, cd(1.0), Rsp, count))) break;
        const double *radii = c->b_pos.as<double>() + 3 * c->N;
        k_mie_pullback<<<(c->N + 63) / 64, 64, 0, c->stream>>>(c->b_mie.as<cd>(), c->N, c->lmax, c->H, c->n, c->nh,
                                                                cols, Y, Rsp, c->k0, radii, gr, ge, gkp);
        if (hipGetLastError() != hipSuccess) {
            set_error("Mie pullback launch failed");
            st = TG_DEVICE_ERROR;
            break;
        }
        std::vector<double> h_pos(3 * c->N), h_r(c->N), h_kp(c->N);
        double h_gk = 0.0;
        std::vector<cd> h_e(c->N);
        hipError_t e1 = hipMemcpyAsync(h_pos.data(), gpos, sizeof(double) * 3 * c->N, hipMemcpyDeviceToHost, c->stream);
        hipError_t e2 = hipMemcpyAsync(h_r.data(), gr, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream);
        hipError_t e3 = hipMemcpyAsync(h_kp.data(), gkp, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream);
        hipError_t e4 = hipMemcpyAsync(&h_gk, gk, sizeof(double), hipMemcpyDeviceToHost, c->stream);
        hipError_t e5 = hipMemcpyAsync(h_e.data(), ge, sizeof(cd) * c->N, hipMemcpyDeviceToHost, c->stream);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess || e5 != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            set_error("pullback copy failed");
            st = TG_DEVICE_ERROR;
            break;
        }
        double k0g = h_gk;
        for (int i = 0; i < c->N; ++i) k0g += h_kp[i];  // sphere order
        if (grad_k0) *grad_k0 = k0g;
        if (grad_radii) memcpy(grad_radii, h_r.data(), sizeof(double) * c->N);
        if (grad_eps)
            for (int i = 0; i < c->N; ++i) grad_eps[i] = tg_c128{h_e[i].re, h_e[i].im};
        if (grad_positions) memcpy(grad_positions, h_pos.data(), sizeof(double) * 3 * c->N);
    } while (0);
    if (conv_host)
        for (int k = 0; k < cols && k < (int)conv.size(); ++k) conv_host[k] = conv[k];
    (void)hipStreamSynchronize(c->stream);
    io.release();
    return st;
}

// Gradients with T an independent input (cluster/interaction.rs:379-440 through the
// expansion pullback of cluster/particles.rs:231-250): Y = A^{-H} G, B-bar = T^H Y,
// the pair gradients of C-bar = T^H Y X^H (positions, and k0 through the coupling),
// and the local blocks T-bar = Y (B + C X)^H.
int pullback_local_with(tg_cluster *c, const tg_c128 *incident, const tg_c128 *x, const tg_c128 *cotangent, int cols,
                        const TangentSolve &adjoint_solve, double *grad_k0, double *grad_positions,
                        tg_c128 *grad_local, tg_c128 *incident_grad, tg_convergence *conv_host) {
    if (!c || !incident || !x || !cotangent || cols < 1) {
        set_error("null argument or no columns");
        return TG_INVALID_INPUT;
    }
    const int64_t count = (int64_t)cols * c->n;
    DevBuf io;
    TG_TRY(io.ensure(sizeof(cd) * count * 6 + sizeof(double) * (3 * (size_t)c->N + 2)));
    cd *a = io.as<cd>(), *X = a + count, *Gs = X + count, *Y = Gs + count, *U = Y + count, *Rsp = U + count;
    double *gpos = reinterpret_cast<double *>(Rsp + count);
    double *gk = gpos + 3 * c->N;
    int st = TG_OK;
    std::vector<tg_convergence> conv;
    do {
        if ((st = ref_to_split(c, (const cd *)incident, a, cols))) break;
        if ((st = ref_to_split(c, (const cd *)x, X, cols))) break;
        if ((st = ref_to_split(c, (const cd *)cotangent, Gs, cols))) break;
        if ((st = adjoint_solve(Gs, Y, cols, conv))) break;
        if ((st = local_apply(c, Y, U, cols, 1, nullptr))) break;  // T^H Y
        if (incident_grad && (st = split_to_ref(c, U, (cd *)incident_grad, cols))) break;
        if ((st = pair_gradients(c, U, X, cols, 0, gpos, gk))) break;
        if (grad_local) {
            if ((st = coupling_matrix_free(c, X, Rsp, cols, 0))) break;
            if ((st = vec_axpby(c->stream, cd(1.0), a, cd(1.0), Rsp, count))) break;
            if ((st = local_cotangent(c, Y, Rsp, cols, (cd *)grad_local))) break;
        }
        std::vector<double> h_pos(3 * c->N);
        double h_gk = 0.0;
        if (hipMemcpyAsync(h_pos.data(), gpos, sizeof(double) * 3 * c->N, hipMemcpyDeviceToHost, c->stream) !=
                hipSuccess ||
            hipMemcpyAsync(&h_gk, gk, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            set_error("pullback copy failed");
            st = TG_DEVICE_ERROR;
            break;
        }
        if (grad_k0) *grad_k0 = h_gk;
        if (grad_positions) memcpy(grad_positions, h_pos.data(), sizeof(double) * 3 * c->N);
    } while (0);
    if (conv_host)
        for (int k = 0; k < cols && k < (int)conv.size(); ++k) conv_host[k] = conv[k];
    (void)hipStreamSynchronize(c->stream);
    io.release();
    return st;
}

}  // namespace tg

extern "C" {

int tg_cluster_pullback_local(tg_cluster *c, const tg_c128 *incident, const tg_c128 *x, const tg_c128 *cotangent,
                              int cols, const tg_gmres_options *o, double *grad_k0, double *grad_positions,
                              tg_c128 *grad_local, tg_c128 *incident_grad, tg_convergence *conv_host) {
    if (!o) {
        set_error("null options");
        return TG_INVALID_INPUT;
    }
    tg_gmres_options od = *o;
    od.atol = 0.0;
    if (!(od.rtol > 0)) od.rtol = 1e-10;
    TangentSolve solve = [c, od](const cd *g, cd *y, int k, std::vector<tg_convergence> &conv) {
        return solve_split(c, g, y, k, od, 1, conv);
    };
    return pullback_local_with(c, incident, x, cotangent, cols, solve, grad_k0, grad_positions, grad_local,
                               incident_grad, conv_host);
}

int tg_cluster_pullback(tg_cluster *c, const tg_c128 *incident, const tg_c128 *x, const tg_c128 *cotangent, int cols,
                        const tg_gmres_options *o, double *grad_k0, double *grad_radii, tg_c128 *grad_eps,
                        double *grad_positions, tg_c128 *incident_grad, tg_convergence *conv_host) {
    if (!o) {
        set_error("null options");
        return TG_INVALID_INPUT;
    }
    tg_gmres_options od = *o;
    od.atol = 0.0;  // derivative solves are relative-only (linalg/gmres.rs:50-60)
    if (!(od.rtol > 0)) od.rtol = 1e-10;
    AdjointSolve solve = [c, od](const cd *g, cd *y, int k, std::vector<tg_convergence> &conv) {
        return solve_split(c, g, y, k, od, 1, conv);
    };
    return pullback_with(c, incident, x, cotangent, cols, solve, grad_k0, grad_radii, grad_eps, grad_positions,
                         incident_grad, conv_host);
}

int tg_cluster_overlap_pullback(tg_cluster *c, const tg_c128 *u, const tg_c128 *v, int cols, double *grad_positions) {
    if (!c || !u || !v || cols < 1) {
        set_error("null argument or no columns");
        return TG_INVALID_INPUT;
    }
    const int64_t count = (int64_t)cols * c->n;
    DevBuf io;
    TG_TRY(io.ensure(sizeof(cd) * count * 2 + sizeof(double) * (3 * (size_t)c->N + 2)));
    cd *us = io.as<cd>(), *vs = us + count;
    double *gpos = reinterpret_cast<double *>(vs + count);
    int st = ref_to_split(c, (const cd *)u, us, cols);
    if (!st) st = ref_to_split(c, (const cd *)v, vs, cols);
    if (!st) st = pair_gradients(c, us, vs, cols, 1, gpos, gpos + 3 * c->N);
    if (!st) {
        hipError_t e = hipMemcpy(grad_positions, gpos, sizeof(double) * 3 * c->N, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            set_error("%s", hipGetErrorString(e));
            st = TG_DEVICE_ERROR;
        }
    }
    io.release();
    return st;
}

}  // extern "C"
