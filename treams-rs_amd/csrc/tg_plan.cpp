// Host-side construction of the helicity translation plan.
//
// For a source mode (l, m) and a destination mode (lambda, mu) of equal helicity the
// translation coefficient is (sw/coupling.rs:86-140, sw/plan.rs:71-170)
//
//   C_pol = sum_p W_same(p) T_p^{q} + (2 pol - 1) sum_p W_cross(p) T_p^{q},  q = m - mu,
//
// with T_p^q = z_p(kd) P_p^q(cos theta) e^{i q phi} and complex weights carrying the
// phase i^{lambda - l + p}.  For `same` degrees p = l + lambda (mod 2) the phase is
// real, for `cross` degrees it is i times a sign, so every weight is stored as one
// real number w and the kernel forms  C_pol = S_same + (2 pol - 1) i S_cross  with
// S = sum w T.  Wigner 3j symbols come from Racah's closed formula evaluated in long
// double (an algorithm independent of the recurrence the CPU oracle restates).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "tg_plan.h"

namespace tg {

static long double factorial_ld(int n) {
    static long double table[200];
    static bool ready = false;
    if (!ready) {
        table[0] = 1.0L;
        for (int i = 1; i < 200; ++i) table[i] = table[i - 1] * (long double)i;
        ready = true;
    }
    return (n >= 0 && n < 200) ? table[n] : 0.0L;
}

double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
    if (j1 < 0 || j2 < 0 || j3 < 0 || m1 + m2 + m3 != 0) return 0.0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
    if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0.0;
    if (j1 + j2 + j3 >= 190) return 0.0;
    long double delta = std::sqrt(factorial_ld(j1 + j2 - j3) * factorial_ld(j1 - j2 + j3) *
                                  factorial_ld(-j1 + j2 + j3) / factorial_ld(j1 + j2 + j3 + 1));
    long double pre = std::sqrt(factorial_ld(j1 + m1) * factorial_ld(j1 - m1) * factorial_ld(j2 + m2) *
                                factorial_ld(j2 - m2) * factorial_ld(j3 + m3) * factorial_ld(j3 - m3));
    int kmin = 0;
    if (j2 - j3 - m1 > kmin) kmin = j2 - j3 - m1;
    if (j1 - j3 + m2 > kmin) kmin = j1 - j3 + m2;
    int kmax = j1 + j2 - j3;
    if (j1 - m1 < kmax) kmax = j1 - m1;
    if (j2 + m2 < kmax) kmax = j2 + m2;
    long double sum = 0.0L;
    for (int k = kmin; k <= kmax; ++k) {
        long double den = factorial_ld(k) * factorial_ld(j1 + j2 - j3 - k) * factorial_ld(j1 - m1 - k) *
                          factorial_ld(j2 + m2 - k) * factorial_ld(j3 - j2 + m1 + k) * factorial_ld(j3 - j1 - m2 + k);
        sum += ((k % 2) ? -1.0L : 1.0L) / den;
    }
    int phase = j1 - j2 - m3;
    long double sign = ((phase % 2) + 2) % 2 ? -1.0L : 1.0L;
    return (double)(sign * delta * pre * sum);
}

// Wigner's small d^l_(m' m)(theta) by his finite sum (z-y-z convention, the element
// <l m'| exp(-i theta J_y) |l m>), in long double:
//   sum_s (-1)^(m' - m + s) sqrt((l+m')! (l-m')! (l+m)! (l-m)!)
//     / ((l+m-s)! s! (m'-m+s)! (l-m'-s)!) cos(theta/2)^(2l+m-m'-2s) sin(theta/2)^(m'-m+2s).
double wigner_small_d(int l, int mp, int m, double theta) {
    if (l < 0 || std::abs(mp) > l || std::abs(m) > l || 2 * l >= 190) return 0.0;
    const long double c = std::cos(0.5L * (long double)theta), s = std::sin(0.5L * (long double)theta);
    const long double pre =
        std::sqrt(factorial_ld(l + mp) * factorial_ld(l - mp) * factorial_ld(l + m) * factorial_ld(l - m));
    int smin = 0, smax = l + m;
    if (m - mp > smin) smin = m - mp;
    if (l - mp < smax) smax = l - mp;
    long double sum = 0.0L;
    for (int k = smin; k <= smax; ++k) {
        const long double den = factorial_ld(l + m - k) * factorial_ld(k) * factorial_ld(mp - m + k) *
                                factorial_ld(l - mp - k);
        const int pc = 2 * l + m - mp - 2 * k, ps = mp - m + 2 * k;
        long double term = pre / den;
        for (int i = 0; i < pc; ++i) term *= c;
        for (int i = 0; i < ps; ++i) term *= s;
        sum += (((mp - m + k) % 2 + 2) % 2 ? -1.0L : 1.0L) * term;
    }
    return (double)sum;
}

// sqrt((p - q)! / (p + q)!)
static double factorial_ratio_sqrt(int p, int q) {
    return (double)std::sqrt(factorial_ld(p - q) / factorial_ld(p + q));
}

int build_plan_rect(int lmax_dst, int lmax_src, HostPlan &plan) {
    if (lmax_dst < 1 || lmax_src < 1 || lmax_dst > 2 * kMaxLmax || lmax_src > 2 * kMaxLmax ||
        lmax_dst + lmax_src > 2 * kMaxLmax)
        return -1;
    const int Hd = lmax_dst * (lmax_dst + 2), Hs = lmax_src * (lmax_src + 2);
    plan.lmax = lmax_dst;
    plan.lmax_src = lmax_src;
    plan.H = Hd;
    plan.Hs = Hs;
    plan.L2 = lmax_dst + lmax_src;
    plan.table_len = (plan.L2 + 1) * (plan.L2 + 1);
    plan.starts.assign((size_t)Hd * Hs + 1, 0);
    plan.nsame.assign((size_t)Hd * Hs, 0);
    plan.w.clear();
    plan.hidx.clear();
    auto modes = [](int lmax, std::vector<int> &ls, std::vector<int> &ms) {
        for (int l = 1; l <= lmax; ++l)
            for (int m = -l; m <= l; ++m) {
                ls.push_back(l);
                ms.push_back(m);
            }
    };
    std::vector<int> lsrc, msrc, ldst, mdst;
    modes(lmax_src, lsrc, msrc);
    modes(lmax_dst, ldst, mdst);
    for (int a = 0; a < Hs; ++a) {       // source (l, m): column
        for (int b = 0; b < Hd; ++b) {   // destination (lambda, mu): row
            const int e = a * Hd + b;
            plan.starts[e] = (int32_t)plan.w.size();
            const int l = lsrc[a], m = msrc[a], lam = ldst[b], mu = mdst[b];
            const int q = m - mu;
            const double pref = 0.5 * ((m % 2 == 0) ? 1.0 : -1.0) *
                                std::sqrt((double)((2 * l + 1) * (2 * lam + 1)) /
                                          ((double)(l * (l + 1)) * (double)(lam * (lam + 1))));
            int count_same = 0;
            for (int cross = 0; cross <= 1; ++cross) {
                const int pmin = std::abs(l - lam) + cross;
                for (int p = l + lam - cross; p >= pmin; p -= 2) {
                    if (p < std::abs(q)) break;
                    double factor;
                    if (cross)
                        factor = std::sqrt((double)(l + lam + 1 + p) * (double)(l + lam + 1 - p) *
                                           (double)(p + lam - l) * (double)(p - lam + l));
                    else
                        factor = (double)(l * (l + 1) + lam * (lam + 1) - p * (p + 1));
                    const double w3a = wigner3j(l, lam, p, m, -mu, mu - m);
                    const double w3b = wigner3j(l, lam, p - cross, 0, 0, 0);
                    double weight = pref * factor * (double)(2 * p + 1) * factorial_ratio_sqrt(p, q) * w3a * w3b;
                    // phase i^{lambda - l + p}: real sign for same, i * sign for cross
                    const int n = lam - l + p - cross;  // even
                    const int half = ((n / 2) % 2 + 2) % 2;
                    if (half) weight = -weight;
                    if (weight == 0.0) continue;
                    plan.w.push_back(weight);
                    plan.hidx.push_back((uint16_t)(p * p + p + q));
                    if (!cross) ++count_same;
                }
            }
            plan.nsame[e] = (uint8_t)count_same;
        }
    }
    plan.starts[(size_t)Hd * Hs] = (int32_t)plan.w.size();
    // reflection factors (-1)^q (p-q)!/(p+q)! for negative orders
    plan.refl.assign(plan.table_len, 0.0);
    for (int p = 0; p <= plan.L2; ++p)
        for (int q = 1; q <= p; ++q)
            plan.refl[p * p + p - q] =
                ((q % 2) ? -1.0 : 1.0) * (double)(factorial_ld(p - q) / factorial_ld(p + q));
    return 0;
}

int build_plan(int lmax, HostPlan &plan) {
    if (lmax < 1 || lmax > kMaxLmax) return -1;
    return build_plan_rect(lmax, lmax, plan);
}

}  // namespace tg
