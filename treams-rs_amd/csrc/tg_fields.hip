// Electric fields of mode expansions at sample points: outgoing (singular) or
// regular helicity vector spherical waves about every sphere of a cluster, or about
// one centre (a global expansion), and the field of a helicity plane wave.
//
// Vector spherical waves (fields.rs:144-320): with the scaled radial function
// c = z_l(k r) / r^l, e = z_{l+1}(k r) / r^{l+1} (z = h for outgoing, j for regular
// waves) and the solid harmonic S = r^l P_l^m(cos theta) e^{i m phi},
//   n = d grad S + l e S r,  d = (l + 1) c / k - r^2 e,     m = c (r x grad S),
//   wave(l, m, pol) = i sqrt((2l+1) / (4 pi l (l+1))) sqrt((l-m)!/(l+m)!) (n - (2 pol - 1) m) / sqrt(2).
// Coordinates are divided by their largest component (the waves do not change
// under a common change of length unit), so r^l stays near one; for regular waves
// the unit is at least 1 / k, and c comes from its series below k r = 1/2 (finite at
// the centre, fields.rs:140-180).
#include <algorithm>
#include <cmath>
#include <vector>

#include "tg_internal.h"
#include "tg_special.h"

namespace tg {

constexpr int kFieldCols = 8;  // coefficient columns per pass

// One thread per (sample point, chunk of spheres): part[chunk][col][3 p + a] =
// sum_{i in chunk} sum_modes x[col][i M + mode] wave_mode(point - pos_i); the chunks
// are summed afterwards in a fixed order.  Per sphere and order |m| the degree
// recurrences of the solid harmonic (value and gradient) and of h_l run side by
// side, so no table indexed by a run-time degree is kept.  The coefficients are read
// with wave-uniform addresses.  A point at a sphere centre gets NaN.
// j_l(x) / r^l for l = 0 .. last (last <= kMaxLmax + 1), x = k r: the series
// k^l / (2l+1)!! sum_q (-x^2 / 2)^q / (q! (2l+3) ... (2l+2q+1)) below x = 1/2 (12
// terms: the tail is below 4e-33 of the first), else j_l(x) / r^l.
__device__ void regular_scaled(int last, double k, double r, double *out) {
    const double x = k * r;
    if (x < 0.5) {
        double kl = 1.0, dfact = 1.0;  // k^l, (2l+1)!!
        for (int l = 0; l <= last; ++l) {
            if (l > 0) {
                kl *= k;
                dfact *= 2.0 * l + 1.0;
            }
            double coef = kl / dfact, sum = 0.0, x2q = 1.0;
            for (int q = 0; q < 12; ++q) {
                sum += coef * x2q;
                x2q *= x * x;
                coef /= -2.0 * (q + 1) * (2.0 * l + 2.0 * q + 3.0);
            }
            out[l] = sum;
        }
        return;
    }
    tg_bessel_j_seq(last, x, out);
    double ir = 1.0;
    for (int l = 0; l <= last; ++l) {
        out[l] *= ir;
        ir /= r;
    }
}

template <bool REG>
__global__ __launch_bounds__(128) void k_cluster_field(int N, int chunk, int lmax, double k, const double *pos,
                                                       const double *refl, const cd *x, int64_t n, int cols,
                                                       int64_t npts, const double *points, cd *part) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const int i0 = blockIdx.y * chunk, i1 = min(i0 + chunk, N);
    const int M = 2 * lmax * (lmax + 2);
    cd E[kFieldCols][3];
#pragma unroll
    for (int c = 0; c < kFieldCols; ++c) E[c][0] = E[c][1] = E[c][2] = cd(0.0);
    bool singular = false;
    for (int i = i0; i < i1; ++i) {
        double rs[3];
        double scale = 0.0;
        for (int a = 0; a < 3; ++a) {
            rs[a] = points[3 * p + a] - pos[3 * i + a];
            scale = fmax(scale, fabs(rs[a]));
        }
        if (REG) scale = fmax(scale, 1.0 / k);
        if (!(scale > 0.0)) {
            singular = true;
            continue;
        }
        for (int a = 0; a < 3; ++a) rs[a] /= scale;
        const double ks = k * scale;
        const double r2 = rs[0] * rs[0] + rs[1] * rs[1] + rs[2] * rs[2], rr = sqrt(r2), irr = 1.0 / rr;
        const double xarg = ks * rr, inv = 1.0 / xarg, sx = sin(xarg), cxv = cos(xarg), iks = 1.0 / ks;
        const cd xy(rs[0], rs[1]);
        const cd *xi = x + (int64_t)i * M;
        double jr[kMaxLmax + 2];  // regular waves: j_l(ks rr) / rr^l
        if (REG) regular_scaled(lmax + 1, ks, rr, jr);
        // ladder over the order: P_q^q (-1)^q (2q-1)!! (x + i y)^q and its gradient
        cd lp(1.0), lg[3] = {cd(0.0), cd(0.0), cd(0.0)};
        for (int q = 0; q <= lmax; ++q) {
            if (q > 0) {
                const double f = -(double)(2 * q - 1);
                const cd e0(1.0), e1(0.0, 1.0);
                lg[0] = f * (lg[0] * xy + lp * e0);
                lg[1] = f * (lg[1] * xy + lp * e1);
                lg[2] = f * (lg[2] * xy);
                lp = f * (lp * xy);
            }
            // degree recurrence at order q; h_l(ks rr) alongside
            cd S = lp, G[3] = {lg[0], lg[1], lg[2]};
            cd Sp(0.0), Gp[3] = {cd(0.0), cd(0.0), cd(0.0)};
            cd hprev(cxv * inv, sx * inv), hcur(sx * inv, -cxv * inv);  // h_{-1}, h_0
            double irpow = 1.0;                                           // rr^-l
            for (int l = 0; l <= lmax; ++l) {
                const cd hnext = ((double)(2 * l + 1) * inv) * hcur - hprev;  // h_{l+1}
                if (l > q) {
                    const double a = (double)(2 * l - 1) / (double)(l - q), b = (double)(l + q - 1) / (double)(l - q);
                    const cd Sn = (a * rs[2]) * S - (b * r2) * Sp;
                    cd Gn[3];
                    for (int c = 0; c < 3; ++c)
                        Gn[c] = a * (rs[2] * G[c] + (c == 2 ? S : cd(0.0))) - b * (r2 * Gp[c] + (2.0 * rs[c]) * Sp);
                    Sp = S;
                    S = Sn;
                    for (int c = 0; c < 3; ++c) {
                        Gp[c] = G[c];
                        G[c] = Gn[c];
                    }
                }
                if (l >= q && l >= 1) {
                    const cd cl = REG ? cd(jr[l]) : irpow * hcur, el = REG ? cd(jr[l + 1]) : (irpow * irr) * hnext;
                    const cd d = ((double)(l + 1) * iks) * cl - r2 * el;
                    const double base = sqrt((2.0 * l + 1.0) / (4.0 * M_PI * l * (l + 1.0)));
                    // (l - q)! / (l + q)!: the plan's reflection factors hold it for q >= 1
                    const double ratio = q == 0 ? 1.0 : fabs(refl[l * l + l - q]);
                    for (int sgn = 0; sgn < (q > 0 ? 2 : 1); ++sgn) {
                        // order m = +q, or m = -q from S^{-q} = (-1)^q (l-q)!/(l+q)! conj(S^q)
                        const int m = sgn ? -q : q;
                        cd Sm = S, Gm[3] = {G[0], G[1], G[2]};
                        double nrm = base * sqrt(ratio);
                        if (sgn) {
                            const double f = (q & 1 ? -1.0 : 1.0) * ratio;
                            Sm = f * conj(S);
                            for (int c = 0; c < 3; ++c) Gm[c] = f * conj(G[c]);
                            nrm = base / sqrt(ratio);
                        }
                        cd nv[3], mv[3];
                        for (int c = 0; c < 3; ++c) nv[c] = d * Gm[c] + ((double)l * rs[c]) * (el * Sm);
                        mv[0] = cl * (rs[1] * Gm[2] - rs[2] * Gm[1]);
                        mv[1] = cl * (rs[2] * Gm[0] - rs[0] * Gm[2]);
                        mv[2] = cl * (rs[0] * Gm[1] - rs[1] * Gm[0]);
                        const cd norm(0.0, nrm * M_SQRT1_2);
                        const int mode = 2 * (l * l + l + m - 1);
                        for (int pol = 1; pol >= 0; --pol) {
                            const double s = pol ? -1.0 : 1.0;  // -(2 pol - 1)
                            cd w[3];
                            for (int c = 0; c < 3; ++c) w[c] = norm * (nv[c] + s * mv[c]);
                            const int64_t row = mode + (1 - pol);
#pragma unroll
                            for (int cc = 0; cc < kFieldCols; ++cc) {
                                if (cc >= cols) break;
                                const cd coef = xi[(int64_t)cc * n + row];
                                for (int c = 0; c < 3; ++c) E[cc][c] += coef * w[c];
                            }
                        }
                    }
                }
                hprev = hcur;
                hcur = hnext;
                irpow *= irr;
            }
        }
    }
    cd *out = part + (int64_t)blockIdx.y * cols * 3 * npts;
#pragma unroll
    for (int cc = 0; cc < kFieldCols; ++cc) {
        if (cc >= cols) break;
        for (int c = 0; c < 3; ++c) out[(int64_t)cc * 3 * npts + 3 * p + c] = singular ? cd(NAN, NAN) : E[cc][c];
    }
}

// out[t] = sum over the chunks of part[chunk][t], in chunk order.
__global__ void k_sum_parts(const cd *part, int parts, int64_t count, cd *out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    cd s(0.0);
    for (int q = 0; q < parts; ++q) s += part[(int64_t)q * count + t];
    out[t] = s;
}

// out[3 p + a] = e_a exp(i k.r_p) for the helicity polarization vector e.
__global__ void k_plane_wave_field(double k0, double d0, double d1, double d2, int pol, int64_t npts,
                                   const double *points, cd *out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    const double kx = k0 * d0, ky = k0 * d1, kz = k0 * d2;
    const double kt = hypot(kx, ky);
    cd Mv[3], Nv[3];
    if (kt == 0.0) {
        Mv[0] = cd(0.0); Mv[1] = cd(0.0, -1.0); Mv[2] = cd(0.0);
        Nv[0] = cd(kz > 0 ? -1.0 : 1.0); Nv[1] = cd(0.0); Nv[2] = cd(0.0);
    } else {
        const double ux = kx / kt, uy = ky / kt;
        Mv[0] = cd(0.0, uy); Mv[1] = cd(0.0, -ux); Mv[2] = cd(0.0);
        Nv[0] = cd(-ux * kz / k0); Nv[1] = cd(-uy * kz / k0); Nv[2] = cd(kt / k0);
    }
    const double s = 2.0 * pol - 1.0;
    const double ph = kx * points[3 * p] + ky * points[3 * p + 1] + kz * points[3 * p + 2];
    const cd phase(cos(ph), sin(ph));
    for (int a = 0; a < 3; ++a) out[3 * p + a] = phase * ((Nv[a] + s * Mv[a]) * M_SQRT1_2);
}

// Fields of N expansions about the centres pos (device, xyz), M modes each, n = N M
// rows per column: (point, centre chunk) threads, chunks summed in order.
int launch_field(int N, int lmax, double k, const double *pos, const double *refl, const cd *x, int64_t n, int cols,
                 int64_t npoints, const double *points, bool regular, cd *out, hipStream_t s) {
    if (npoints == 0) return TG_OK;
    // enough (point, centre chunk) threads to fill the chip: about 64 waves per CU
    const int64_t target = (int64_t)256 * 64 * 64;
    int parts = (int)std::min<int64_t>(N, std::max<int64_t>(1, target / std::max<int64_t>(npoints, 1)));
    parts = std::min(parts, 64);
    const int chunk = (N + parts - 1) / parts;
    parts = (N + chunk - 1) / chunk;
    DevBuf scratch;
    TG_TRY(scratch.ensure(sizeof(cd) * (size_t)parts * kFieldCols * 3 * npoints));
    for (int c0 = 0; c0 < cols; c0 += kFieldCols) {
        const int nc = std::min(kFieldCols, cols - c0);
        cd *dst = out + (int64_t)c0 * 3 * npoints;
        const dim3 grid((unsigned)((npoints + 127) / 128), (unsigned)parts);
        cd *target_buf = parts == 1 ? dst : scratch.as<cd>();
        if (regular)
            k_cluster_field<true><<<grid, 128, 0, s>>>(N, chunk, lmax, k, pos, refl, x + (int64_t)c0 * n, n, nc,
                                                       npoints, points, target_buf);
        else
            k_cluster_field<false><<<grid, 128, 0, s>>>(N, chunk, lmax, k, pos, refl, x + (int64_t)c0 * n, n, nc,
                                                        npoints, points, target_buf);
        TG_HIP(hipGetLastError());
        if (parts > 1) {
            const int64_t count = (int64_t)nc * 3 * npoints;
            k_sum_parts<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(scratch.as<cd>(), parts, count, dst);
            TG_HIP(hipGetLastError());
        }
    }
    TG_HIP(hipStreamSynchronize(s));
    scratch.release();
    return TG_OK;
}

}  // namespace tg

using namespace tg;

extern "C" {

int tg_cluster_field(tg_cluster *c, const tg_c128 *x, int cols, int64_t npoints, const double *points, int regular,
                     tg_c128 *out) {
    if (!c || !x || !points || !out || cols < 1 || npoints < 0) {
        set_error("null argument, no columns or negative point count");
        return TG_INVALID_INPUT;
    }
    return launch_field(c->N, c->lmax, c->k0, c->b_pos.as<double>(), c->plan.refl, (const cd *)x, c->n, cols,
                        npoints, points, regular != 0, (cd *)out, c->stream);
}

int tg_plane_wave_field(double k0, const double *direction, int pol, int64_t npoints, const double *points,
                        tg_c128 *out, void *stream) {
    if (!direction || !(std::isfinite(k0) && k0 > 0) || (pol != 0 && pol != 1) || npoints < 0 ||
        (npoints > 0 && (!points || !out))) {
        set_error("invalid wavenumber, direction, polarization or points");
        return TG_INVALID_INPUT;
    }
    const double nd = std::sqrt(direction[0] * direction[0] + direction[1] * direction[1] +
                                direction[2] * direction[2]);
    if (!(nd > 0) || !std::isfinite(nd)) {
        set_error("plane wave needs a finite nonzero direction");
        return TG_INVALID_INPUT;
    }
    if (npoints == 0) return TG_OK;
    hipStream_t s = (hipStream_t)stream;
    k_plane_wave_field<<<(unsigned)((npoints + 255) / 256), 256, 0, s>>>(
        k0, direction[0] / nd, direction[1] / nd, direction[2] / nd, pol, npoints, points, (cd *)out);
    TG_HIP(hipGetLastError());
    TG_HIP(hipStreamSynchronize(s));
    return TG_OK;
}

}  // extern "C"
