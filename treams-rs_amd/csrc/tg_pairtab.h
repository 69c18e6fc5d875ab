// Singular harmonic tables of sphere pairs, h_p(k r) P_p^q(cos theta) e^{i q phi}
// (sw/plan.rs:194-289 evaluates the same products per block), shared by the FP64 pair
// tables (tg_kernels.hip) and the split-FP16 tables built straight from the
// recurrences (tg_split16.hip), so both see the same values.
#pragma once
#include "tg_complex.h"

namespace tg {

// Length and unit direction of d, scaling by the largest component first;
// returns 0 where every component is below DBL_MIN (the origin).
__device__ __forceinline__ int unit_direction(const double d[3], double &r, double u[3]) {
    double s = fmax(fabs(d[0]), fmax(fabs(d[1]), fabs(d[2])));
    if (s < 2.2250738585072014e-308) return 0;
    double a = d[0] / s, b = d[1] / s, c = d[2] / s;
    double nrm = sqrt(a * a + b * b + c * c);
    r = s * nrm;
    u[0] = a / nrm;
    u[1] = b / nrm;
    u[2] = c / nrm;
    return 1;
}

// The recurrence steps with explicit FMAs, so that every kernel that evaluates the
// tables rounds alike whatever the compiler would contract around them.
__device__ __forceinline__ cd cmul_fma(cd a, cd b) {
    return cd(fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re));
}
// h_{p+1} = (2p + 1) / x h_p - h_{p-1}, with s = (2p + 1) / x
__device__ __forceinline__ cd hankel_next(double s, cd hcur, cd hprev) {
    return cd(fma(s, hcur.re, -hprev.re), fma(s, hcur.im, -hprev.im));
}
// S_p = a z S_{p-1} - b S_{p-2}, with az = a z
__device__ __forceinline__ cd legendre_next(double az, double b, cd cur, cd prev) {
    return cd(fma(az, cur.re, -(b * prev.re)), fma(az, cur.im, -(b * prev.im)));
}
// the reflected (negative-order) harmonic refl conj(S)
__device__ __forceinline__ cd reflect(double refl, cd cur) { return cd(refl * cur.re, -(refl * cur.im)); }

// The singular table one order q at a time: the diagonal ladder P_q^q = (-1)^q
// (2q-1)!! (x + i y)^q carried from order to order, the degree recurrence
// (p - q) S_p = (2p - 1) z S_{p-1} - (p + q - 1) S_{p-2}, and the Hankel upward
// recurrence run alongside each order from h_{-1}, h_0 (an array indexed by a run-time
// degree would live in scratch memory).  Negative orders by S_p^{-q} = (-1)^q
// (p-q)!/(p+q)! conj(S_p^q) (refl).  put(h, value) stores harmonic index h.
struct SingularState {
    double inv, sx, cx;
    cd xy, ladder;
};

__device__ __forceinline__ void singular_begin(double x, const double u[3], SingularState &s) {
    s.inv = 1.0 / x;
    s.sx = sin(x);
    s.cx = cos(x);
    s.xy = cd(u[0], u[1]);
    s.ladder = cd(1.0);
}

// orders must come in ascending sequence q = 0, 1, ... (the ladder advances here)
template <class Put>
__device__ __forceinline__ void singular_column_put(int q, int L2, const double u[3], const double *refl,
                                                    SingularState &s, Put put) {
    if (q > 0) s.ladder = (-(double)(2 * q - 1)) * cmul_fma(s.ladder, s.xy);
    cd hprev(s.cx * s.inv, s.sx * s.inv), hcur(s.sx * s.inv, -s.cx * s.inv);  // h_{-1}, h_0
    cd cur = s.ladder, prev(0.0);
    for (int p = 0; p <= L2; ++p) {
        if (p > q) {
            const double a = (double)(2 * p - 1) / (double)(p - q);
            const double b = (double)(p + q - 1) / (double)(p - q);
            const cd next = legendre_next(a * u[2], b, cur, prev);
            prev = cur;
            cur = next;
        }
        if (p >= q) {
            put(p * p + p + q, cmul_fma(hcur, cur));
            if (q > 0) put(p * p + p - q, cmul_fma(hcur, reflect(refl[p * p + p - q], cur)));
        }
        const cd hnext = hankel_next((double)(2 * p + 1) * s.inv, hcur, hprev);
        hprev = hcur;
        hcur = hnext;
    }
}

template <class Put>
__device__ __forceinline__ void singular_table_put(int L2, double x, const double u[3], const double *refl,
                                                   Put put) {
    SingularState s;
    singular_begin(x, u, s);
    for (int q = 0; q <= L2; ++q) singular_column_put(q, L2, u, refl, s, put);
}

// exponent e with max(|re|, |im|) < 2^e (frexp), or -32768 for zero
__device__ __forceinline__ int entry_exponent(cd v) {
    const double m = fmax(fabs(v.re), fabs(v.im));
    int e = 0;
    (void)frexp(m, &e);
    return m > 0.0 ? e : -32768;
}

}  // namespace tg
