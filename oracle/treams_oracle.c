/*
 * FP64 CPU oracle for the cluster-coupling path (test infrastructure only).
 *
 * This file restates, in plain C, the reference algorithms of treams-rs that the
 * GPU library replaces.  It is never linked into the product library: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it, as the
 * checker.  Every function names the reference file:line it follows (paths relative
 * to crates/treams-core/src/ of the reference).
 *
 * Conventions (sw/mod.rs:17-31): a translation from a source to a destination uses
 * the displacement destination - source; pol 1 (positive helicity) precedes pol 0;
 * blocks are column-major with the source mode outer and the destination inner.
 *
 * Parity pinning (what checks each part) is listed in tests/test_oracle_golden.py.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------- factorials -- */

/* ln n!, special/mod.rs:104-115 (libm lgamma). */
static double log_factorial(int n) { return lgamma((double)n + 1.0); }

/* sqrt((l-m)!/(l+m)!), special/mod.rs:118-120. */
static double factorial_ratio_sqrt(int l, int m) {
    return exp(0.5 * (log_factorial(l - m) - log_factorial(l + m)));
}

static double parity(int n) { return (n % 2 == 0) ? 1.0 : -1.0; }

/* ----------------------------------------------------------------- wigner 3j -- */
/* Two-sided Schulten-Gordon recurrence, special/wigner.rs:36-317. */

static double w3j_c(int j1, int j2, int j3, int m3) { /* wigner.rs:36-41 */
    double a = (double)j3 * j3 - (double)(j1 - j2) * (j1 - j2);
    double b = (double)(j1 + j2 + 1) * (j1 + j2 + 1) - (double)j3 * j3;
    double c = (double)j3 * j3 - (double)m3 * m3;
    return sqrt(a * b * c);
}

static double w3j_d(int j1, int j2, int j3, int m1, int m2, int m3) { /* wigner.rs:45-49 */
    return (double)(2 * j3 + 1) *
           ((double)(j3 * (j3 + 1)) * (double)(m2 - m1) + (double)(j2 * (j2 + 1) - j1 * (j1 + 1)) * (double)m3);
}

static double w3j_initial_j(int j1, int j2, int m1, int m2) { /* wigner.rs:53-66 */
    return parity(j1 + m1) *
           exp(0.5 * (log_factorial(j1 - m1) + log_factorial(j1 + m1) + log_factorial(2 * j1 - 2 * j2) +
                      log_factorial(2 * j2) - log_factorial(j2 - m2) - log_factorial(j2 + m2) -
                      log_factorial(j1 - j2 - m1 - m2) - log_factorial(j1 - j2 + m1 + m2) -
                      log_factorial(2 * j1 + 1)));
}

static double w3j_initial_m(int j1, int j2, int m1, int m2) { /* wigner.rs:70-83 */
    return parity(j2 + m2) *
           exp(0.5 * (log_factorial(j1 + m1) + log_factorial(j2 + m2) + log_factorial(j1 + j2 - m1 - m2) +
                      log_factorial(2 * m1 + 2 * m2) - log_factorial(j1 - m1) - log_factorial(j2 - m2) -
                      log_factorial(j1 - j2 + m1 + m2) - log_factorial(j2 - j1 + m1 + m2) -
                      log_factorial(j1 + j2 + m1 + m2 + 1)));
}

static double w3j_lowest(int j1, int j2, int m1, int m2, int minimum) { /* wigner.rs:87-97 */
    if (minimum == j1 - j2) return w3j_initial_j(j1, j2, m1, m2);
    if (minimum == j2 - j1) return w3j_initial_j(j2, j1, m2, m1);
    if (minimum == m1 + m2) return w3j_initial_m(j1, j2, m1, m2);
    return w3j_initial_m(j2, j1, -m2, -m1);
}

static double w3j_highest(int j1, int j2, int m1, int m2) { /* wigner.rs:101-115 */
    int maximum = j1 + j2, m3 = -m1 - m2;
    return parity(maximum - m3) *
           exp(0.5 * (log_factorial(2 * j1) + log_factorial(2 * j2) + log_factorial(maximum + m3) +
                      log_factorial(maximum - m3) - log_factorial(2 * maximum + 1) - log_factorial(j1 - m1) -
                      log_factorial(j1 + m1) - log_factorial(j2 - m2) - log_factorial(j2 + m2)));
}

static int w3j_growth_end(int j1, int j2, int m1, int m2, int minimum, int stop) { /* wigner.rs:120-135 */
    int m3 = -m1 - m2;
    double ratio = INFINITY;
    for (int j = minimum + 1; j <= stop; ++j) {
        if (minimum == 0 && j == 1)
            ratio = -(double)(m2 - m1) / w3j_c(j1, j2, 1, m3);
        else
            ratio = (-w3j_d(j1, j2, j - 1, m1, m2, m3) - (double)j * w3j_c(j1, j2, j - 1, m3) / ratio) /
                    ((double)(j - 1) * w3j_c(j1, j2, j, m3));
        if (fabs(ratio) < 1.0) return j;
    }
    return stop + 1;
}

static int w3j_decay_start(int j1, int j2, int m1, int m2, int maximum, int stop) { /* wigner.rs:140-155 */
    int m3 = -m1 - m2;
    double ratio = INFINITY;
    for (int j = maximum - 1; j >= stop; --j) {
        if (j == maximum - 1)
            ratio = -w3j_d(j1, j2, maximum, m1, m2, m3) / ((double)(maximum + 1) * w3j_c(j1, j2, maximum, m3));
        else
            ratio = (-w3j_d(j1, j2, j + 1, m1, m2, m3) - (double)(j + 1) * w3j_c(j1, j2, j + 2, m3) / ratio) /
                    ((double)(j + 2) * w3j_c(j1, j2, j + 1, m3));
        if (fabs(ratio) < 1.0) return j + 1;
    }
    return stop;
}

static int w3j_turn(int j1, int j2, int m1, int m2, int minimum, int maximum, int first, int last) {
    /* wigner.rs:161-180 */
    int switch_at = (j1 + j2 - abs(j1 - j2)) / 4 + abs(j1 - j2);
    int growth = w3j_growth_end(j1, j2, m1, m2, minimum, last);
    if (growth > last) return last + 1;
    int bound = growth > first ? growth : first;
    if (switch_at <= bound) return bound;
    int decay = w3j_decay_start(j1, j2, m1, m2, maximum, bound);
    return switch_at < decay ? switch_at : decay;
}

/* Row (j1 j2 j3; m1 m2 -m1-m2) for j3 in first..=j1+j2 into out[j3-first]; wigner.rs:199-252, 292-307. */
static int w3j_row(int j1, int j2, int m1, int m2, double *out, int *first_out) {
    if (j1 < 0 || j2 < 0 || abs(m1) > j1 || abs(m2) > j2) { *first_out = 0; return 0; }
    int m3 = -m1 - m2;
    int first = abs(j1 - j2) > abs(m1 + m2) ? abs(j1 - j2) : abs(m1 + m2);
    int last = j1 + j2;
    int minimum = abs(j1 - j2) > abs(m3) ? abs(j1 - j2) : abs(m3);
    int maximum = j1 + j2;
    *first_out = first;
    int count = last + 1 - first;
    if (count <= 0) return 0;
    for (int i = 0; i < count; ++i) out[i] = 0.0;
    int turn = w3j_turn(j1, j2, m1, m2, minimum, maximum, first, last);
    if (first < turn) {
        double value = w3j_lowest(j1, j2, m1, m2, minimum), previous = 0.0;
        int stop = turn < last + 1 ? turn : last + 1;
        for (int j = minimum; j < stop; ++j) {
            if (j > minimum) {
                double next;
                if (minimum == 0 && j == 1)
                    next = -value * (double)(m2 - m1) / w3j_c(j1, j2, 1, m3);
                else
                    next = (-w3j_d(j1, j2, j - 1, m1, m2, m3) * value -
                            (double)j * w3j_c(j1, j2, j - 1, m3) * previous) /
                           ((double)(j - 1) * w3j_c(j1, j2, j, m3));
                previous = value;
                value = next;
            }
            if (j >= first) out[j - first] = value;
        }
    }
    if (turn > last) return count;
    double value = w3j_highest(j1, j2, m1, m2), previous = 0.0;
    if (last == maximum) out[maximum - first] = value;
    int lo = turn > first ? turn : first;
    for (int j = maximum - 1; j >= lo; --j) {
        double next;
        if (j == maximum - 1)
            next = -w3j_d(j1, j2, j + 1, m1, m2, m3) * value / ((double)(j + 2) * w3j_c(j1, j2, j + 1, m3));
        else
            next = (-w3j_d(j1, j2, j + 1, m1, m2, m3) * value -
                    (double)(j + 1) * w3j_c(j1, j2, j + 2, m3) * previous) /
                   ((double)(j + 2) * w3j_c(j1, j2, j + 1, m3));
        previous = value;
        value = next;
        if (j <= last) out[j - first] = value;
    }
    return count;
}

/* Single symbol (j1 j2 j3; m1 m2 m3); wigner.rs:263-279 (via the row). */
EXPORT double or_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (j1 < 0 || j2 < 0 || j3 < 0 || j3 < abs(j1 - j2) || j3 > j1 + j2 || abs(m1) > j1 || abs(m2) > j2 ||
        abs(m3) > j3 || m1 + m2 + m3 != 0)
        return 0.0;
    double row[600];
    int first;
    int count = w3j_row(j1, j2, m1, m2, row, &first);
    if (j3 - first < 0 || j3 - first >= count) return 0.0;
    return row[j3 - first];
}

static double row_get(const double *row, int first, int count, int j3) {
    int i = j3 - first;
    return (i >= 0 && i < count) ? row[i] : 0.0;
}

/* --------------------------------------------------------------- mode order -- */

/* sw/mod.rs:114-125: l ascending, m from -l to l, pol 1 before pol 0. */
EXPORT int or_modes(int lmax, int *l_out, int *m_out, int *pol_out) {
    int k = 0;
    for (int l = 1; l <= lmax; ++l)
        for (int m = -l; m <= l; ++m)
            for (int pol = 1; pol >= 0; --pol) {
                if (l_out) { l_out[k] = l; m_out[k] = m; pol_out[k] = pol; }
                ++k;
            }
    return k;
}

/* Admitted degrees of sw/coupling.rs:146-160, descending in steps of two. */
EXPORT int or_degrees(int l, int m, int lambda, int mu, int cross, int *out) {
    int start = l + lambda - cross;
    int end = abs(lambda - l) + cross;
    if (abs(m - mu) > end) end = abs(m - mu);
    int n = (start - end + 2) / 2;
    if (start - end + 2 < 0) n = 0;
    for (int k = 0; k < n; ++k)
        if (out) out[k] = start - 2 * k;
    return n;
}

/* ------------------------------------------------------------ coupling terms -- */

/* i^n for integer n. */
static cplx ipow(int n) {
    int r = ((n % 4) + 4) % 4;
    return r == 0 ? 1.0 : r == 1 ? I : r == 2 ? -1.0 : -I;
}

/* sw/coupling.rs:19-41 with mu' = -mu passed by the caller. */
static cplx tl_vsw_term(int l, int m, int lambda, int mu, int p, int q, const double *row0, int f0, int c0,
                        const double *row1, int f1, int c1) {
    int lo = abs(m + mu) > abs(l - lambda) ? abs(m + mu) : abs(l - lambda);
    if (p < lo || p > l + lambda || q < abs(l - lambda) || q > l + lambda || ((q + l + lambda) % 2) != 0)
        return 0.0;
    return (double)(2 * p + 1) * ipow(lambda - l + p) * factorial_ratio_sqrt(p, m + mu) *
           row_get(row0, f0, c0, p) * row_get(row1, f1, c1, q);
}

/*
 * Terms of one (source (l,m) -> destination (lambda,mu)) coupling in the helicity
 * basis, sw/coupling.rs:86-140 with Kinds::of (62-68): equal helicities get the
 * `same` terms, then the `cross` terms multiplied by 2 pol - 1.  Opposite
 * helicities get none.  Writes (p, weight) and returns the count.
 */
static int coupling_terms(int l, int m, int lambda, int mu, int pol_to, int pol_from, int *p_out, cplx *w_out) {
    if (pol_to != pol_from) return 0;
    double row0[64], row1[64];
    int f0, f1;
    int c0 = w3j_row(l, lambda, m, -mu, row0, &f0);
    int c1 = w3j_row(l, lambda, 0, 0, row1, &f1);
    double sign = (m % 2 == 0) ? 1.0 : -1.0;
    double pref = 0.5 * sign *
                  sqrt((double)((2 * l + 1) * (2 * lambda + 1)) /
                       ((double)(l * (l + 1)) * (double)(lambda * (lambda + 1))));
    double hsign = 2.0 * pol_from - 1.0; /* special/polarization.rs:141-143 */
    int count = 0;
    for (int cross = 0; cross <= 1; ++cross) {
        int ps[64];
        int n = or_degrees(l, m, lambda, mu, cross, ps);
        for (int k = 0; k < n; ++k) {
            int p = ps[k];
            double factor;
            if (cross)
                factor = sqrt((double)(l + lambda + 1 + p) * (double)(l + lambda + 1 - p) * (double)(p + lambda - l) *
                              (double)(p - lambda + l));
            else
                factor = (double)(l * (l + 1) + lambda * (lambda + 1) - p * (p + 1));
            cplx w = pref * factor * tl_vsw_term(l, m, lambda, -mu, p, p - cross, row0, f0, c0, row1, f1, c1);
            if (creal(w) * creal(w) + cimag(w) * cimag(w) > 0.0) {
                p_out[count] = p;
                w_out[count] = cross ? hsign * w : w;
                ++count;
            }
        }
    }
    return count;
}

/*
 * CSR translation plan between all modes up to lmax in the helicity basis,
 * sw/plan.rs:71-170: entry e = source * M + destination owns
 * terms[starts[e]..starts[e+1]] of {index = p^2 + p + (m - mu), weight}.
 * With index == NULL only the total count is returned (starts still filled).
 */
EXPORT int or_plan(int lmax, int64_t *starts, int32_t *index, cplx *weight) {
    int M = or_modes(lmax, NULL, NULL, NULL);
    int *L = malloc(sizeof(int) * M), *Mm = malloc(sizeof(int) * M), *P = malloc(sizeof(int) * M);
    or_modes(lmax, L, Mm, P);
    int64_t total = 0;
    int ps[64];
    cplx ws[64];
    for (int s = 0; s < M; ++s)
        for (int d = 0; d < M; ++d) {
            if (starts) starts[s * M + d] = total;
            int n = coupling_terms(L[s], Mm[s], L[d], Mm[d], P[d], P[s], ps, ws);
            if (index)
                for (int k = 0; k < n; ++k) {
                    index[total + k] = ps[k] * ps[k] + ps[k] + Mm[s] - Mm[d];
                    weight[total + k] = ws[k];
                }
            total += n;
        }
    if (starts) starts[M * M] = total;
    free(L); free(Mm); free(P);
    return (int)total;
}

/* ---------------------------------------------------------- radial functions -- */

/*
 * Outgoing spherical Hankel h_p(x), p = 0..last, at real positive x: the seed
 * h_-1 = e^{ix}/x, h_0 = -i e^{ix}/x and upward recurrence (special/bessel.rs:495-516).
 * The reference qualifies this for x in [0.5, 1e4]; below 0.5 it calls AMOS, which
 * the oracle does not have, so the same recurrence is used there and pinned by
 * 50-digit mpmath fixtures instead.
 */
EXPORT void or_hankels(int last, double x, cplx *out) {
    double inv = 1.0 / x, s = sin(x), c = cos(x);
    cplx previous = c * inv + I * (s * inv);
    cplx current = s * inv - I * (c * inv);
    for (int p = 0; p <= last; ++p) {
        out[p] = current;
        if (p < last) {
            cplx next = ((double)(2 * p + 1) * inv) * current - previous;
            previous = current;
            current = next;
        }
    }
}

/*
 * Regular spherical Bessel j_p(z), p = 0..last, complex z.  Below |z| = 0.5 the
 * 16-term even power series of special/bessel.rs:425-431; elsewhere (AMOS in the
 * reference, un-vendored complex-bessel 0.3.0) a backward ratio recurrence
 * r_p = j_p / j_{p-1} = z / (2p + 1 - z r_{p+1}) normalised by the larger of
 * j_0 = sin z / z and j_1.  Pinned by mpmath fixtures, including the zeros of j_0
 * and j_1 (z = pi, 2 pi, 3 pi, 4.4934...).
 */
static void sph_j_series(int l, cplx z, cplx *value, cplx *first, cplx *second);

EXPORT void or_spherical_jn(int last, cplx z, cplx *out) {
    if (cabs(z) < 0.5) {
        for (int p = 0; p <= last; ++p) {
            cplx f, d1, d2;
            sph_j_series(p, z, &f, &d1, &d2);
            out[p] = f;
        }
        return;
    }
    int start = last + (int)cabs(z) + 60;
    cplx r = 0.0;
    cplx *ratios = malloc(sizeof(cplx) * (last + 2));
    for (int p = start; p >= 1; --p) {
        r = z / ((double)(2 * p + 1) - z * r);
        if (p <= last) ratios[p] = r;
    }
    /* one scale for the whole sequence, from whichever of j_0 = sin z / z and
     * j_1 = (sin z / z - cos z) / z is larger: their zeros interlace, so the larger
     * is never near a zero (scaling by j_0 alone loses j_1.. near z = k pi) */
    cplx j0 = csin(z) / z, j1 = (j0 - ccos(z)) / z;
    out[0] = j0;
    if (last >= 1) out[1] = cabs(j0) >= cabs(j1) ? j0 * ratios[1] : j1;
    for (int p = 2; p <= last; ++p) out[p] = out[p - 1] * ratios[p];
    free(ratios);
}

/* The even series of special/bessel.rs:378-406 for j_l (425-431), with derivatives. */
static void sph_j_series(int l, cplx z, cplx *value, cplx *first, cplx *second) {
    double dfact = 1.0;
    for (int i = 0; i <= l; ++i) dfact *= (double)(2 * i + 1);
    double leading = 1.0 / dfact;
    cplx tv = leading * cpow(z, l);
    cplx tf = l > 0 ? leading * (double)l * cpow(z, l - 1) : 0.0;
    cplx ts = l > 1 ? leading * (double)(l * (l - 1)) * cpow(z, l - 2) : 0.0;
    cplx sv = tv, sf = tf, ss = ts, sq = z * z;
    for (int q = 1; q < 16; ++q) {
        double factor = -1.0 / (2.0 * q * (2.0 * (l + q) + 1.0));
        cplx nv = factor * sq * tv;
        cplx nf = factor * (sq * tf + 2.0 * z * tv);
        cplx ns = factor * (sq * ts + 4.0 * z * tf + 2.0 * tv);
        tv = nv; tf = nf; ts = ns;
        sv += tv; sf += tf; ss += ts;
    }
    *value = sv; *first = sf; *second = ss;
}

/* Scaled complex division, numerics/mod.rs:63-66. */
static cplx ratio(cplx a, cplx b) {
    double s = fabs(creal(b)) > fabs(cimag(b)) ? fabs(creal(b)) : fabs(cimag(b));
    return (a / s) / (b / s);
}

/* Jet (value, first, second) from the next order, special/bessel.rs:152-160. */
static void jet_from_next(int l, cplx z, cplx value, cplx next, cplx *jet) {
    double order = (double)l;
    cplx first = order * ratio(value, z) - next;
    cplx second = -2.0 * ratio(first, z) + order * (order + 1.0) * ratio(ratio(value, z), z) - value;
    jet[0] = value; jet[1] = first; jet[2] = second;
}

/*
 * Radial jets of degrees 0..last at z: regular (j) or singular (outgoing h, real z
 * only), special/bessel.rs:453-487 (series below |z| = 0.5 for regular).
 * jets has 3 * (last + 1) entries.
 */
EXPORT void or_radial_jets(int last, cplx z, int regular, cplx *jets) {
    if (regular && cabs(z) < 0.5) {
        for (int p = 0; p <= last; ++p) sph_j_series(p, z, &jets[3 * p], &jets[3 * p + 1], &jets[3 * p + 2]);
        return;
    }
    cplx *values = malloc(sizeof(cplx) * (last + 2));
    if (regular)
        or_spherical_jn(last + 1, z, values);
    else
        or_hankels(last + 1, creal(z), values);
    for (int p = 0; p <= last; ++p) jet_from_next(p, z, values[p], values[p + 1], &jets[3 * p]);
    free(values);
}

/* ----------------------------------------------------------- solid harmonics -- */

/*
 * r^p P_p^m(z/r) e^{i m phi} for p <= order, |m| <= p at `position`, with the
 * Cartesian gradient when grad != NULL: special/harmonics.rs:147-305 (SolidTable,
 * visit_impl without Hessian), Condon-Shortley phase, negative orders by the
 * reflection factor (-1)^m (p-m)!/(p+m)! times the conjugate.  Entry (p, m) at
 * p^2 + p + m.
 */
EXPORT void or_solid_table(int order, const double *position, cplx *value, cplx *grad) {
    double x = position[0], y = position[1], z = position[2];
    double r2 = x * x + y * y + z * z;
    cplx xy = x + I * y;
    cplx e[3] = {1.0, I, 0.0};
    cplx diag = 1.0, dgrad[3] = {0, 0, 0};
    for (int mm = 0; mm <= order; ++mm) {
        if (mm > 0) {
            double factor = -(double)(2 * mm - 1);
            for (int i = 0; i < 3; ++i) dgrad[i] = factor * (dgrad[i] * xy + diag * e[i]);
            diag *= factor * xy;
        }
        cplx p = diag, g[3] = {dgrad[0], dgrad[1], dgrad[2]};
        cplx prev = 0.0, pg[3] = {0, 0, 0};
        for (int deg = mm; deg <= order; ++deg) {
            if (deg > mm) {
                double a = (double)(2 * deg - 1) / (double)(deg - mm);
                double b = (double)(deg + mm - 1) / (double)(deg - mm);
                cplx next = a * z * p - b * r2 * prev;
                cplx ng[3];
                for (int i = 0; i < 3; ++i)
                    ng[i] = a * (z * g[i] + (i == 2 ? p : 0.0)) - b * (r2 * pg[i] + 2.0 * position[i] * prev);
                prev = p;
                for (int i = 0; i < 3; ++i) pg[i] = g[i];
                p = next;
                for (int i = 0; i < 3; ++i) g[i] = ng[i];
            }
            int idx = deg * deg + deg + mm;
            value[idx] = p;
            if (grad)
                for (int i = 0; i < 3; ++i) grad[3 * idx + i] = g[i];
            if (mm > 0) {
                int nidx = deg * deg + deg - mm;
                double refl = parity(mm) * exp(log_factorial(deg - mm) - log_factorial(deg + mm));
                value[nidx] = refl * conj(p);
                if (grad)
                    for (int i = 0; i < 3; ++i) grad[3 * nidx + i] = refl * conj(g[i]);
            }
        }
    }
}

/* Length and unit direction, special/harmonics.rs:108-116; returns 0 at the origin. */
static int direction(const double *position, double *r, double *unit) {
    double scale = 0.0;
    for (int i = 0; i < 3; ++i)
        if (fabs(position[i]) > scale) scale = fabs(position[i]);
    if (scale < DBL_MIN) return 0;
    double s[3], norm = 0.0;
    for (int i = 0; i < 3; ++i) { s[i] = position[i] / scale; norm += s[i] * s[i]; }
    norm = sqrt(norm);
    *r = scale * norm;
    for (int i = 0; i < 3; ++i) unit[i] = s[i] / norm;
    return 1;
}

/* ---------------------------------------------------------------- the plan -- */

typedef struct {
    int lmax, M, order, len;
    int64_t *starts;
    int32_t *index;
    cplx *weight;
} Plan;

static Plan plan_new(int lmax) {
    Plan p;
    p.lmax = lmax;
    p.M = or_modes(lmax, NULL, NULL, NULL);
    p.order = 2 * lmax;
    p.len = (p.order + 1) * (p.order + 1);
    p.starts = malloc(sizeof(int64_t) * (p.M * p.M + 1));
    int total = or_plan(lmax, p.starts, NULL, NULL);
    p.index = malloc(sizeof(int32_t) * (total > 0 ? total : 1));
    p.weight = malloc(sizeof(cplx) * (total > 0 ? total : 1));
    or_plan(lmax, p.starts, p.index, p.weight);
    return p;
}

static void plan_free(Plan *p) { free(p->starts); free(p->index); free(p->weight); }

/* Harmonic table z_p(k r) P_p^m e^{i m phi}, sw/plan.rs:224-251. */
static void plan_values(const Plan *plan, double k, const double *d, int regular, cplx *table) {
    for (int i = 0; i < plan->len; ++i) table[i] = 0.0;
    double r, unit[3];
    if (!direction(d, &r, unit)) {
        if (regular) table[0] = 1.0; /* j_0(0) = 1 */
        return;
    }
    cplx *solid = malloc(sizeof(cplx) * plan->len);
    or_solid_table(plan->order, unit, solid, NULL);
    cplx *rad = malloc(sizeof(cplx) * (plan->order + 2));
    if (regular)
        or_spherical_jn(plan->order, k * r, rad);
    else
        or_hankels(plan->order, k * r, rad);
    for (int p = 0; p <= plan->order; ++p)
        for (int m = -p; m <= p; ++m) {
            int idx = p * p + p + m;
            table[idx] = rad[p] * solid[idx];
        }
    free(solid);
    free(rad);
}

/* Block = sum weight * table[index], column-major, sw/plan.rs:194-203. */
static void plan_evaluate_table(const Plan *plan, const cplx *table, cplx *block) {
    int MM = plan->M * plan->M;
    for (int e = 0; e < MM; ++e) {
        cplx s = 0.0;
        for (int64_t t = plan->starts[e]; t < plan->starts[e + 1]; ++t) s += plan->weight[t] * table[plan->index[t]];
        block[e] = s;
    }
}

/* One M x M translation block at displacement d (destination - source), sw/plan.rs:282-289. */
EXPORT void or_translation_block(int lmax, double k, const double *d, int regular, cplx *block) {
    Plan plan = plan_new(lmax);
    cplx *table = malloc(sizeof(cplx) * plan.len);
    plan_values(&plan, k, d, regular, table);
    plan_evaluate_table(&plan, table, block);
    free(table);
    plan_free(&plan);
}

/*
 * Pull a block cotangent back to the displacement gradient and the complex
 * wavenumber cotangent, sw/plan.rs:596-615 with the derivative table of
 * plan.rs:254-278 and sw/cartesian.rs:113-128 (away from the origin).
 */
static void plan_pullback(const Plan *plan, const cplx *table_solid, const cplx *table_grad, const cplx *jets,
                          double k, double r, const double *unit, const cplx *cot, double *grad, cplx *gk) {
    int MM = plan->M * plan->M;
    cplx *contracted = calloc(plan->len, sizeof(cplx));
    for (int e = 0; e < MM; ++e) {
        cplx g = cot[e];
        if (g == 0.0) continue;
        for (int64_t t = plan->starts[e]; t < plan->starts[e + 1]; ++t)
            contracted[plan->index[t]] += conj(plan->weight[t]) * g;
    }
    grad[0] = grad[1] = grad[2] = 0.0;
    *gk = 0.0;
    for (int p = 0; p <= plan->order; ++p) {
        cplx value_r = jets[3 * p], first_r = jets[3 * p + 1];
        for (int m = -p; m <= p; ++m) {
            int idx = p * p + p + m;
            cplx Y = table_solid[idx];
            cplx tangent[3];
            for (int a = 0; a < 3; ++a) tangent[a] = table_grad[3 * idx + a] - (double)p * Y * unit[a];
            cplx dk = first_r * r * Y;
            cplx scaled = value_r / r;
            *gk += contracted[idx] * conj(dk);
            for (int a = 0; a < 3; ++a) {
                cplx dpos = first_r * k * unit[a] * Y + scaled * tangent[a];
                grad[a] += creal(conj(contracted[idx]) * dpos);
            }
        }
    }
    free(contracted);
}

EXPORT int or_translation_pullback(int lmax, double k, const double *d, int regular, const cplx *cot, double *grad,
                                   cplx *gk) {
    Plan plan = plan_new(lmax);
    double r, unit[3];
    if (!direction(d, &r, unit)) { plan_free(&plan); return -1; }
    cplx *solid = malloc(sizeof(cplx) * plan.len), *sgrad = malloc(sizeof(cplx) * 3 * plan.len);
    or_solid_table(plan.order, unit, solid, sgrad);
    cplx *jets = malloc(sizeof(cplx) * 3 * (plan.order + 1));
    or_radial_jets(plan.order, k * r, regular, jets);
    plan_pullback(&plan, solid, sgrad, jets, k, r, unit, cot, grad, gk);
    free(solid); free(sgrad); free(jets);
    plan_free(&plan);
    return 0;
}

/*
 * Dense N*M x N*M matrix of pair blocks (column-major), block (i, j) at displacement
 * pos_i - pos_j: cluster/spheres.rs:174-177 and 188-220 (singular C, self blocks
 * exactly zero) or sw/expansion.rs:116-130, 391-398 (regular R, self blocks = the
 * regular plan at the origin = identity).
 */
EXPORT void or_pair_matrix(int lmax, double k0, int N, const double *positions, int regular, cplx *out) {
    Plan plan = plan_new(lmax);
    int M = plan.M;
    int64_t n = (int64_t)N * M;
    cplx *table = malloc(sizeof(cplx) * plan.len), *block = malloc(sizeof(cplx) * M * M);
    for (int64_t i = 0; i < n * n; ++i) out[i] = 0.0;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            if (i == j && !regular) continue;
            double d[3];
            for (int a = 0; a < 3; ++a) d[a] = positions[3 * i + a] - positions[3 * j + a];
            plan_values(&plan, k0, d, regular, table);
            plan_evaluate_table(&plan, table, block);
            for (int c = 0; c < M; ++c)
                memcpy(&out[((int64_t)j * M + c) * n + (int64_t)i * M], &block[c * M], sizeof(cplx) * M);
        }
    free(table);
    free(block);
    plan_free(&plan);
}

/*
 * Position and wavenumber gradients of a dense pair matrix from its cotangent Gbar
 * (n x n, column-major): per pair the block cotangent goes through the plan pullback
 * at pos_i - pos_j, adding +g to particle i and -g to particle j
 * (cluster/iterative.rs:500-551; sw/expansion.rs pullback for R).
 */
EXPORT void or_pair_matrix_pullback(int lmax, double k0, int N, const double *positions, int regular,
                                    const cplx *gbar, double *grad_pos, double *gk0) {
    Plan plan = plan_new(lmax);
    int M = plan.M;
    int64_t n = (int64_t)N * M;
    cplx *solid = malloc(sizeof(cplx) * plan.len), *sgrad = malloc(sizeof(cplx) * 3 * plan.len);
    cplx *jets = malloc(sizeof(cplx) * 3 * (plan.order + 1));
    cplx *cot = malloc(sizeof(cplx) * M * M);
    for (int i = 0; i < 3 * N; ++i) grad_pos[i] = 0.0;
    double total_k = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            if (i == j) continue;
            double d[3], r, unit[3];
            for (int a = 0; a < 3; ++a) d[a] = positions[3 * i + a] - positions[3 * j + a];
            if (!direction(d, &r, unit)) continue;
            for (int c = 0; c < M; ++c)
                for (int rr = 0; rr < M; ++rr) cot[c * M + rr] = gbar[((int64_t)j * M + c) * n + (int64_t)i * M + rr];
            or_solid_table(plan.order, unit, solid, sgrad);
            or_radial_jets(plan.order, k0 * r, regular, jets);
            double g[3];
            cplx gk;
            plan_pullback(&plan, solid, sgrad, jets, k0, r, unit, cot, g, &gk);
            for (int a = 0; a < 3; ++a) { grad_pos[3 * i + a] += g[a]; grad_pos[3 * j + a] -= g[a]; }
            total_k += creal(gk);
        }
    *gk0 = total_k;
    free(solid); free(sgrad); free(jets); free(cot);
    plan_free(&plan);
}

/*
 * Matrix-free y_i = sum_{j != i} C_ij x_j for `cols` column-major vectors, the
 * per-pair plan application of cluster/iterative.rs:187-221 with sw/plan.rs:327-360
 * (harmonic table once per pair, shared by the columns).  With regular != 0 the
 * regular overlap R instead (j_p radial functions, identity self blocks:
 * sw/expansion.rs:116-130 with Radial::Regular), the R x of the cross sections
 * (python/treams_rs/_tmatrix.py:404-411).  Only destination particles in [i0, i1)
 * are computed (a bounded sample for timing); OpenMP over destinations, each thread
 * owning its rows.
 */
EXPORT void or_pair_apply(int lmax, double k0, int N, const double *positions, int cols, const cplx *x, cplx *y,
                          int i0, int i1, int regular) {
    Plan plan = plan_new(lmax);
    int M = plan.M;
    int64_t n = (int64_t)N * M;
#pragma omp parallel
    {
        cplx *table = malloc(sizeof(cplx) * plan.len);
#pragma omp for schedule(dynamic, 1)
        for (int i = i0; i < i1; ++i) {
            for (int c = 0; c < cols; ++c)
                for (int r = 0; r < M; ++r)
                    y[(int64_t)c * n + (int64_t)i * M + r] = regular ? x[(int64_t)c * n + (int64_t)i * M + r] : 0.0;
            for (int j = 0; j < N; ++j) {
                if (j == i) continue;
                double d[3];
                for (int a = 0; a < 3; ++a) d[a] = positions[3 * i + a] - positions[3 * j + a];
                plan_values(&plan, k0, d, regular, table);
                for (int s = 0; s < M; ++s)
                    for (int dd = 0; dd < M; ++dd) {
                        int e = s * M + dd;
                        if (plan.starts[e] == plan.starts[e + 1]) continue;
                        cplx v = 0.0;
                        for (int64_t t = plan.starts[e]; t < plan.starts[e + 1]; ++t)
                            v += plan.weight[t] * table[plan.index[t]];
                        for (int c = 0; c < cols; ++c)
                            y[(int64_t)c * n + (int64_t)i * M + dd] += v * x[(int64_t)c * n + (int64_t)j * M + s];
                    }
            }
        }
        free(table);
    }
    plan_free(&plan);
}

/* Matrix-free C x, cluster/iterative.rs:187-221 (self blocks zero). */
EXPORT void or_coupling_apply(int lmax, double k0, int N, const double *positions, int cols, const cplx *x, cplx *y,
                              int i0, int i1) {
    or_pair_apply(lmax, k0, N, positions, cols, x, y, i0, i1, 0);
}

/* --------------------------------------------------------------------- Mie -- */

/* Riccati entries (f, f' + f/x) and their x-derivatives, coeffs/mie.rs:23-37. */
typedef struct { cplx wave[2][2]; cplx slope[2][2]; } Riccati; /* [kind][0: value, 1: f' + f/x] */

static void riccati(int l, cplx x, int need_singular, Riccati *out) {
    cplx jets[3 * 130];
    for (int kind = 0; kind < 2; ++kind) {
        if (kind == 1 && !need_singular) {
            out->wave[1][0] = out->wave[1][1] = out->slope[1][0] = out->slope[1][1] = 0.0;
            continue;
        }
        or_radial_jets(l, x, kind == 0, jets);
        cplx f = jets[3 * l], f1 = jets[3 * l + 1], f2 = jets[3 * l + 2];
        out->wave[kind][0] = f;
        out->wave[kind][1] = f1 + f / x;
        out->slope[kind][0] = f1;
        out->slope[kind][1] = f2 + f1 / x - f / (x * x);
    }
}

/*
 * Interface matrix F of coeffs/mie.rs:134-181 for one boundary between an achiral
 * inner medium (index n, impedance z0) and vacuum, restricted to the columns of
 * the regular inner waves (the only ones the core q of mie.rs:217-219 selects).
 * With tangent != 0 also its directional derivative along (dx_inner, dz_inner).
 */
typedef struct { cplx f[4][2]; } Interface;

static void interface_matrix(cplx x_in, cplx x_out, cplx z_in, cplx z_out, const Riccati *in, const Riccati *out,
                             cplx dx_in, cplx dx_out, cplx dz_in, int tangent, Interface *F, Interface *dF) {
    cplx ratio_z = z_out / z_in;
    cplx da = (0.0 / z_in - ratio_z * dz_in / z_in) / (2.0 * I);
    for (int p = 0; p < 2; ++p) {
        cplx square = x_out * x_out;
        for (int q = 0; q < 2; ++q) {
            double sign = (p == q) ? 1.0 : -1.0;
            cplx a = (ratio_z + sign) / (2.0 * I);
            cplx factor = a * square;
            cplx dfactor = tangent ? (da * square + 2.0 * a * x_out * dx_out) : 0.0;
            int kind = 0; /* regular inner waves only */
            cplx inner0 = in->wave[kind][0], inner1 = in->wave[kind][1];
            cplx dinner0 = in->slope[kind][0] * dx_in, dinner1 = in->slope[kind][1] * dx_in;
            for (int out_kind = 0; out_kind < 2; ++out_kind) {
                int ok = 1 - out_kind;
                cplx outer0 = out->wave[ok][0], outer1 = out->wave[ok][1];
                cplx douter0 = out->slope[ok][0] * dx_out, douter1 = out->slope[ok][1] * dx_out;
                double out_sign = out_kind == 0 ? 1.0 : -1.0;
                cplx term = out_sign * (outer1 * inner0 - sign * outer0 * inner1);
                F->f[p + 2 * out_kind][q] = term * factor;
                if (tangent) {
                    cplx dterm = out_sign * (douter1 * inner0 + outer1 * dinner0 - sign * (douter0 * inner1 + outer0 * dinner1));
                    dF->f[p + 2 * out_kind][q] = dterm * factor + term * dfactor;
                }
            }
        }
    }
}

/* Scaled 2x2 inverse, coeffs/mie.rs:186-201; row-major a[2][2]. */
static int inverse2(cplx a[2][2], cplx out[2][2]) {
    double scale = 0.0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            double v = fabs(creal(a[i][j])) > fabs(cimag(a[i][j])) ? fabs(creal(a[i][j])) : fabs(cimag(a[i][j]));
            if (v > scale) scale = v;
        }
    if (scale == 0.0 || !isfinite(scale)) return -1;
    cplx b[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) b[i][j] = a[i][j] / scale;
    cplx det = b[0][0] * b[1][1] - b[0][1] * b[1][0];
    if (det == 0.0) return -1;
    out[0][0] = ratio(b[1][1], det) / scale;
    out[0][1] = ratio(-b[0][1], det) / scale;
    out[1][0] = ratio(-b[1][0], det) / scale;
    out[1][1] = ratio(b[0][0], det) / scale;
    return 0;
}

typedef struct {
    cplx x_in, x_out, z_in, z_out, n;
    Riccati in, out;
    Interface F;
    cplx inv[2][2];  /* inverse of the upper rows */
    cplx value[2][2]; /* helicity order (negative, positive) */
} MieState;

/* Mie matrix of degree l of a homogeneous sphere in vacuum, coeffs/mie.rs:209-255. */
static int mie_state(int l, double size, cplx eps, MieState *s) {
    s->n = csqrt(eps);
    s->x_in = size * s->n;
    s->x_out = size;
    s->z_in = csqrt(1.0 / eps);
    s->z_out = 1.0;
    riccati(l, s->x_in, 0, &s->in);
    riccati(l, s->x_out, 1, &s->out);
    interface_matrix(s->x_in, s->x_out, s->z_in, s->z_out, &s->in, &s->out, 0, 0, 0, 0, &s->F, NULL);
    cplx upper[2][2] = {{s->F.f[0][0], s->F.f[0][1]}, {s->F.f[1][0], s->F.f[1][1]}};
    if (inverse2(upper, s->inv)) return -1;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) s->value[i][j] = s->F.f[2 + i][0] * s->inv[0][j] + s->F.f[2 + i][1] * s->inv[1][j];
    return 0;
}

/* 2x2 Mie matrix in helicity order (row-major, index 0 = negative). */
EXPORT int or_mie(int l, double size, cplx eps, cplx *out) {
    MieState s;
    if (mie_state(l, size, eps, &s)) return -1;
    out[0] = s.value[0][0]; out[1] = s.value[0][1]; out[2] = s.value[1][0]; out[3] = s.value[1][1];
    return 0;
}

/*
 * Gradients of a real loss with cotangent G (helicity order, row-major) with respect
 * to the size parameter and the inner permittivity: coeffs/mie.rs:306-349 for one
 * boundary (gq rows from the cotangent, contraction conj(sum conj(gf) dF)).
 */
EXPORT int or_mie_pullback(int l, double size, cplx eps, const cplx *cot, double *g_size, cplx *g_eps) {
    MieState s;
    if (mie_state(l, size, eps, &s)) return -1;
    cplx G[2][2] = {{cot[0], cot[1]}, {cot[2], cot[3]}};
    /* lower = G inverse^H ; upper = -value^H lower */
    cplx lower[2][2], upper[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) lower[i][j] = G[i][0] * conj(s.inv[j][0]) + G[i][1] * conj(s.inv[j][1]);
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) upper[i][j] = -(conj(s.value[0][i]) * lower[0][j] + conj(s.value[1][i]) * lower[1][j]);
    /* gq (4x2); before = core selector, so gf = gq restricted to the regular columns */
    cplx gf[4][2];
    for (int j = 0; j < 2; ++j) { gf[0][j] = upper[0][j]; gf[1][j] = upper[1][j]; gf[2][j] = lower[0][j]; gf[3][j] = lower[1][j]; }
    Interface F, dF;
    /* size direction: dx = x / size on both sides */
    interface_matrix(s.x_in, s.x_out, s.z_in, s.z_out, &s.in, &s.out, s.x_in / size, s.x_out / size, 0, 1, &F, &dF);
    cplx acc = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 2; ++j) acc += conj(gf[i][j]) * dF.f[i][j];
    *g_size = creal(conj(acc));
    /* epsilon direction of the inner medium: dn = mu / (2n), dz = -z / (2 eps) (material.rs:73-97) */
    cplx dn = 1.0 / (2.0 * s.n), dz = -s.z_in / (2.0 * eps);
    interface_matrix(s.x_in, s.x_out, s.z_in, s.z_out, &s.in, &s.out, size * dn, 0.0, dz, 1, &F, &dF);
    acc = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 2; ++j) acc += conj(gf[i][j]) * dF.f[i][j];
    *g_eps = conj(acc);
    return 0;
}

/* ----------------------------------------------------------- plane waves -- */

/* F_l^m = P_l^m / sin^|m| for degrees |m|..max, special/legendre.rs:26-60. */
static void legendre_factors(int max, int m, double z, double *out /* indexed by degree */) {
    int order = abs(m);
    if (max < order) return;
    double p = 1.0;
    if (m >= 0)
        for (int k = 1; k <= m; ++k) p *= -(double)(2 * k - 1);
    else
        for (int k = 1; k <= -m; ++k) p *= 1.0 / (double)(2 * k);
    out[order] = p;
    if (max == order) return;
    double prev = p;
    p *= ((double)(2 * order + 1) / (double)(order - m + 1)) * z;
    out[order + 1] = p;
    for (int k = order + 2; k <= max; ++k) {
        double next = ((double)(2 * k - 1) * z * p - (double)(k + m - 1) * prev) * (1.0 / (double)(k - m));
        prev = p;
        p = next;
        out[k] = p;
    }
}

/* tau + (2 pol - 1) pi, special/legendre.rs:82-122 (helicity). */
static double polarized_angular(int l, int m, double cosine, double sine, int pol) {
    double F[140];
    double pi_v, tau_v;
    if (m == 0) {
        legendre_factors(l, 1, cosine, F);
        pi_v = 0.0;
        tau_v = sine * F[l];
    } else {
        legendre_factors(l, m, cosine, F);
        double upper = F[l], lower = (l > abs(m)) ? F[l - 1] : 0.0;
        double power = pow(sine, abs(m) - 1);
        pi_v = (double)m * power * upper;
        tau_v = power * ((double)l * cosine * upper - (double)(l + m) * lower);
    }
    return tau_v + (2.0 * pol - 1.0) * pi_v;
}

/*
 * Regular local expansion of a unit-amplitude plane wave of helicity `pol` with
 * real unit direction `dir` and vacuum wavenumber k0, at N positions, in mode order:
 * pw/expand.rs:23-66 (prefactor 2 sqrt(pi (2l+1)/(l(l+1))) sqrt((l-m)!/(l+m)!) i^l,
 * azimuth e^{-i m phi}, angular tau + (2pol-1) pi) times the position phase
 * e^{i k.r} of pw/field.rs:20-26.
 */
EXPORT void or_plane_wave(int lmax, int N, const double *positions, double k0, const double *dir, int pol, cplx *out) {
    double norm = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    double kv[3] = {k0 * dir[0] / norm, k0 * dir[1] / norm, k0 * dir[2] / norm};
    double k = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
    double transverse = sqrt(kv[0] * kv[0] + kv[1] * kv[1]);
    int axis = transverse == 0.0;
    double xy[2] = {axis ? 0.0 : kv[0] / transverse, axis ? 0.0 : kv[1] / transverse};
    double cosine = axis ? (kv[2] / k >= 0.0 ? 1.0 : -1.0) : kv[2] / k;
    double sine = transverse / k;
    int M = or_modes(lmax, NULL, NULL, NULL);
    int *L = malloc(sizeof(int) * M), *Mm = malloc(sizeof(int) * M), *P = malloc(sizeof(int) * M);
    or_modes(lmax, L, Mm, P);
    for (int j = 0; j < N; ++j) {
        double angle = kv[0] * positions[3 * j] + kv[1] * positions[3 * j + 1] + kv[2] * positions[3 * j + 2];
        cplx phase = cos(angle) + I * sin(angle);
        for (int e = 0; e < M; ++e) {
            cplx v = 0.0;
            if (P[e] == pol) {
                int l = L[e], m = Mm[e];
                cplx pref = 2.0 * sqrt(M_PI * (2 * l + 1) / (double)(l * (l + 1))) * factorial_ratio_sqrt(l, m) * ipow(l);
                cplx azimuth = axis ? 1.0 : cpow(xy[0] - I * xy[1], m);
                v = pref * azimuth * polarized_angular(l, m, cosine, sine, pol) * phase;
            }
            out[(int64_t)j * M + e] = v;
        }
    }
    free(L); free(Mm); free(P);
}
