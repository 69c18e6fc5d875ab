// Special functions of the cluster path, FP64, host and device.
//
//  - outgoing spherical Hankel h_p(x), real x: seed h_-1 = e^{ix}/x, h_0 = -i e^{ix}/x
//    and upward recurrence (stable: h is the dominant solution);
//  - regular spherical Bessel j_p(x), real x: Miller's backward ratio recurrence
//    normalised by j_0 = sin x / x, power series below x = 0.5;
//  - Mie coefficients of a homogeneous sphere in vacuum by the Bohren-Huffman
//    logarithmic-derivative algorithm (Bohren & Huffman 1983, sec. 4.8), carried out
//    on dual numbers so the same code gives dT/dx and dT/d(epsilon).
#pragma once
#include "tg_complex.h"

// h_p(x), p = 0..last.
TG_HD void tg_hankel_seq(int last, double x, cd *out) {
    double inv = 1.0 / x, s = sin(x), c = cos(x);
    cd prev(c * inv, s * inv);
    cd cur(s * inv, -c * inv);
    for (int p = 0; p <= last; ++p) {
        out[p] = cur;
        cd next = ((double)(2 * p + 1) * inv) * cur - prev;
        prev = cur;
        cur = next;
    }
}

// j_p(x) by its power series x^p/(2p+1)!! sum_q (-x^2/2)^q / (q! (2p+3)...(2p+2q+1)).
TG_HD double tg_sph_j_series(int p, double x) {
    double lead = 1.0;
    for (int i = 1; i <= p; ++i) lead *= x / (double)(2 * i + 1);
    double term = lead, sum = lead, x2 = -0.5 * x * x;
    for (int q = 1; q < 18; ++q) {
        term *= x2 / ((double)q * (double)(2 * p + 2 * q + 1));
        sum += term;
    }
    return sum;
}

// j_p(x), p = 0..last (last <= 63), real x >= 0.  The ratios r_p = j_p / j_{p-1} of
// the backward recurrence fix the sequence up to one factor, taken from whichever of
// j_0 = sin x / x and j_1 = (sin x / x - cos x) / x is larger in magnitude: their
// zeros interlace, so the larger one is never near a zero (at x = pi, j_0 = 0 while
// j_1 = 1 / pi), and every j_p keeps its full relative accuracy.
TG_HD void tg_bessel_j_seq(int last, double x, double *out) {
    if (x < 0.5) {
        for (int p = 0; p <= last; ++p) out[p] = tg_sph_j_series(p, x);
        return;
    }
    int start = last + (int)x + 40;
    double r = 0.0;
    double ratios[72];
    for (int p = start; p >= 1; --p) {
        r = x / ((double)(2 * p + 1) - x * r);
        if (p <= last) ratios[p] = r;
    }
    const double s = sin(x), j0 = s / x, j1 = (j0 - cos(x)) / x;
    out[0] = j0;
    if (last < 1) return;
    out[1] = fabs(j0) >= fabs(j1) ? j0 * ratios[1] : j1;
    for (int p = 2; p <= last; ++p) out[p] = out[p - 1] * ratios[p];
}

// ------------------------------------------------------------------ dual numbers --
// Value and derivatives along the size parameter x and the permittivity epsilon.
struct tg_dual {
    cd v, dx, de;
    TG_HD tg_dual() {}
    TG_HD tg_dual(cd a) : v(a) {}
    TG_HD tg_dual(cd a, cd b, cd c) : v(a), dx(b), de(c) {}
};
TG_HD tg_dual operator+(tg_dual a, tg_dual b) { return {a.v + b.v, a.dx + b.dx, a.de + b.de}; }
TG_HD tg_dual operator-(tg_dual a, tg_dual b) { return {a.v - b.v, a.dx - b.dx, a.de - b.de}; }
TG_HD tg_dual operator-(tg_dual a) { return {-a.v, -a.dx, -a.de}; }
TG_HD tg_dual operator*(tg_dual a, tg_dual b) {
    return {a.v * b.v, a.dx * b.v + a.v * b.dx, a.de * b.v + a.v * b.de};
}
TG_HD tg_dual operator*(double s, tg_dual a) { return {s * a.v, s * a.dx, s * a.de}; }
// 1 / z with one real division: z scaled exactly by the power of two 2^-e that brings
// its larger component into [0.5, 1), so |z 2^-e|^2 lies in [0.25, 2] whatever |z|.
TG_HD cd crecip_scaled(cd z) {
    int e;
    (void)frexp(fmax(fabs(z.re), fabs(z.im)), &e);
    const double a = ldexp(z.re, -e), b = ldexp(z.im, -e);
    const double r = 1.0 / (a * a + b * b);
    return cd(ldexp(a * r, -e), ldexp(-b * r, -e));
}
// One complex reciprocal (one real division) for the value and both derivatives (the
// recurrences below are chains of these divisions).
TG_HD tg_dual operator/(tg_dual a, tg_dual b) {
    const cd r = crecip_scaled(b.v);
    const cd q = a.v * r;
    return {q, (a.dx - q * b.dx) * r, (a.de - q * b.de) * r};
}
TG_HD tg_dual operator/(double s, tg_dual b) { return tg_dual(cd(s)) / b; }

// Mie coefficients a_l, b_l of one degree l (1 <= l <= lmax <= 40) of a nonmagnetic
// sphere of size x and relative permittivity eps in vacuum, with derivatives along x
// and eps (Bohren and Huffman's recurrences; the starting degrees depend on lmax).
// Each degree is computed alone, without arrays:
// * D_l(mx), the logarithmic derivative, downward from zero at degree
//   max(lmax, |mx|) + 16;
// * psi_n(x) = x j_n(x) as psi_0 (or psi_1 in closed form) times the ratios
//   psi_n / psi_{n-1}, themselves downward from zero at degree lmax + x + 40;
// * xi_n(x) = x h_n(x) upward from xi_0, xi_-1.
// Returns 0, or -1 for a zero denominator.
TG_HD int tg_mie_bh(int l, int lmax, double xv, cd eps, tg_dual &a, tg_dual &b) {
    cd mv = csqrt_(eps);
    tg_dual x(cd(xv), cd(1.0), cd(0.0));
    tg_dual m(mv, cd(0.0), cd(0.5) / mv);  // dm/deps = 1/(2m)
    tg_dual mx = m * x;
    double amx = cabs_(mx.v);
    int nstart = (lmax > (int)amx ? lmax : (int)amx) + 16;
    tg_dual d(cd(0.0));
    for (int n = nstart; n > l; --n) {
        tg_dual nz = (double)n / mx;
        d = nz - 1.0 / (d + nz);
    }
    const tg_dual Dl = d;
    // ratios r_n = psi_n / psi_{n-1}; products of r_2 .. r_l and r_2 .. r_{l-1}
    int jstart = lmax + (int)xv + 40;
    tg_dual r(cd(0.0)), q_l(cd(1.0), cd(0.0), cd(0.0)), q_lm(cd(1.0), cd(0.0), cd(0.0)), r1;
    for (int n = jstart; n >= 1; --n) {
        r = x / (tg_dual(cd((double)(2 * n + 1))) - x * r);
        if (n >= 2 && n <= l) q_l = q_l * r;
        if (n >= 2 && n <= l - 1) q_lm = q_lm * r;
        if (n == 1) r1 = r;
    }
    double s = sin(xv), c = cos(xv);
    tg_dual psi0(cd(s), cd(c), cd(0.0));                   // psi_0 = sin x
    tg_dual xi_prev(cd(s, -c), cd(c, s), cd(0.0));         // xi_0 = -i e^{ix}
    tg_dual xi_prev2(cd(c, s), cd(-s, c), cd(0.0));        // xi_-1 = e^{ix}
    // psi_1 = sin x / x - cos x in closed form where it exceeds psi_0 (near x = k pi
    // the ratio chain would scale a vanishing psi_0)
    const double p1 = s / xv - c;
    const bool psi1_closed = fabs(p1) > fabs(s);
    const tg_dual psi1 = psi1_closed ? tg_dual(cd(p1), cd(c / xv - s / (xv * xv) + s), cd(0.0)) : psi0 * r1;
    const tg_dual psi = psi1 * q_l, psi_prev = l == 1 ? psi0 : psi1 * q_lm;
    for (int n = 1; n < l; ++n) {
        tg_dual xi = ((double)(2 * n - 1) / x) * xi_prev - xi_prev2;
        xi_prev2 = xi_prev;
        xi_prev = xi;
    }
    tg_dual xi = ((double)(2 * l - 1) / x) * xi_prev - xi_prev2;
    tg_dual nx = (double)l / x;
    tg_dual ga = Dl / m + nx, gb = m * Dl + nx;
    tg_dual da = ga * xi - xi_prev, db = gb * xi - xi_prev;
    if (abs2(da.v) == 0.0 || abs2(db.v) == 0.0) return -1;
    a = (ga * psi - psi_prev) / da;
    b = (gb * psi - psi_prev) / db;
    return 0;
}
