#pragma once
#include <cstdint>
#include <vector>

namespace tg {

constexpr int kMaxLmax = 12;  // harmonic degree 2 lmax <= 24, table length <= 625

struct HostPlan {
    int lmax = 0, H = 0, L2 = 0, table_len = 0;  // lmax, H: destination side
    int lmax_src = 0, Hs = 0;                      // source side (= lmax, H when square)
    std::vector<int32_t> starts;   // H*Hs + 1, entry e = a * H + b (source a outer)
    std::vector<uint8_t> nsame;    // same-kind terms first in each entry
    std::vector<double> w;         // real weights (cross ones multiply i)
    std::vector<uint16_t> hidx;    // harmonic index p^2 + p + q
    std::vector<double> refl;      // (-1)^q (p-q)!/(p+q)! at index p^2+p-q
};

double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3);
// d^l_(m' m)(theta), z-y-z convention (special/wigner.rs:411)
double wigner_small_d(int l, int mp, int m, double theta);
int build_plan(int lmax, HostPlan &plan);
// Destination degrees up to lmax_dst, source degrees up to lmax_src (harmonic degree
// up to their sum, at most 2 kMaxLmax).
int build_plan_rect(int lmax_dst, int lmax_src, HostPlan &plan);

}  // namespace tg
