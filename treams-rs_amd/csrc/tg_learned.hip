// Learned near-field preconditioner (no reference counterpart; right preconditioning,
// so it changes iteration counts and never the answer, which the FP64 true residual
// certifies).
//
//   M^{-1} x = x + T ( sum_{(i,j) in K-nearest pairs} D(alpha_ij) C_ij x_j )
//
// D(alpha) scales the rows of destination degree l by alpha_ij[l] (complex, both
// helicities, all orders).  alpha = 1 is the first-order Neumann preconditioner of the
// near field; an MLP maps per-pair features (k d, direction cosines, the two spheres'
// Mie blocks) to alpha - 1.  Training minimises E_V |A M^{-1} V - V|^2 / |V|^2 over
// random probe blocks V, whose gradient with respect to alpha_ij[l] is
//   2 sum_{rows r of degree l of sphere i} Z_i[r] conj((C_ij V_j)[r]) / |V|^2,
// Z = T^H A^H (A M^{-1} V - V).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "tg_gmres.h"
#include "tg_internal.h"
#include "tg_mlp.h"
#include "tg_precond.h"

struct LearnedState {
    int K = 0;
    int64_t npairs = 0;
    tg::DevBuf meta;    // pi[np], pj[np], row_start[N+1], col_start[N+1], col_list[np]
    tg::DevBuf disp;      // [np][3] pos_i - pos_j of the near pairs
    tg::DevBuf blocks;    // [np][M x M] c128, reference layout, column-major (built on first use)
    bool have_blocks = false;
    tg::DevBuf blocksh;   // [np][2][H][H] bfloat16-pair same-helicity blocks (forward applies of mixed-precision solves)
    bool have_blocksh = false;
    tg::DevBuf part;      // forward apply: weighted pair products [np][4][M]
    tg::DevBuf feats;   // [np][64] fp32
    tg::DevBuf out16;   // [np][16] fp32 network output
    tg::DevBuf alpha;   // [np][lmax] c128
    tg::DevBuf work;    // training scratch
    const int *pi = nullptr, *pj = nullptr, *row_start = nullptr, *col_start = nullptr, *col_list = nullptr;
    bool have_cols = false;  // col_start and col_list filled
    bool have_alpha = false;
};

namespace tg {

constexpr int kFeat = 64, kOut = 16;

static inline int deg_of_row(int r) { return (int)std::sqrt((double)((r >> 1) + 1) + 0.5); }

__device__ __forceinline__ int row_degree(int r) { return (int)sqrt((double)((r >> 1) + 1) + 0.5); }

// features: k d, 1/(k d), d/|d| (d = pos_i - pos_j), then the 2x2 Mie blocks (diagonal
// and off-diagonal, re/im) of sphere i and of sphere j for l = 1..6; zero padded.
__global__ void k_pair_features(int64_t np, const int *pi, const int *pj, const double *pos, const cd *mie, int lmax,
                                double k0, float *out) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int i = pi[p], j = pj[p];
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = pos[3 * i + a] - pos[3 * j + a];
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    float *f = out + p * kFeat;
    for (int t = 0; t < kFeat; ++t) f[t] = 0.f;
    f[0] = (float)(k0 * r);
    f[1] = (float)(1.0 / (k0 * r));
    for (int a = 0; a < 3; ++a) f[2 + a] = (float)(d[a] / r);
    for (int s = 0; s < 2; ++s) {
        const int q = s == 0 ? i : j;
        for (int l = 0; l < lmax && l < 6; ++l) {
            const cd *b = mie + (size_t)(q * lmax + l) * 4;
            float *g = f + 5 + s * 24 + l * 4;
            g[0] = (float)b[0].re; g[1] = (float)b[0].im; g[2] = (float)b[1].re; g[3] = (float)b[1].im;
        }
    }
}

// alpha[l] = 1 + 0.5 tanh(out[2l]) + 0.5 i tanh(out[2l+1]): bounded around the
// first-order Neumann weights alpha = 1 (out = 0).
__global__ void k_alpha_from_output(int64_t np, int lmax, const float *out, cd *alpha) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np * lmax) return;
    const int64_t p = t / lmax;
    const int l = (int)(t - p * lmax);
    const float re = l < kOut / 2 ? out[p * kOut + 2 * l] : 0.f, im = l < kOut / 2 ? out[p * kOut + 2 * l + 1] : 0.f;
    alpha[t] = cd(1.0 + 0.5 * tanh((double)re), 0.5 * tanh((double)im));
}

__device__ __forceinline__ cd block_value(cd v) { return v; }
__device__ __forceinline__ cd block_value(cf v) { return cd((double)v.re, (double)v.im); }

// y_i = x_i + T_i sum_{pairs p of i} D(alpha_p) C_p x_{j_p} (split layout), in two
// steps.  k_near_pair_products: one wave per near pair p (many short independent
// streams keep enough block loads in flight), lane l the block rows l + 64 k (k < RS,
// reference mode order), the source values x_j wave-uniform (scalar loads);
// part[p][cc][r] = alpha_p(l_r) (C_p x_j)[r] for up to 4 columns.  k_near_finish: one
// workgroup per sphere sums its pairs' parts in pair order (deterministic), applies
// T_i and adds x_i.
template <int RS, class B>
__global__ __launch_bounds__(64) void k_near_pair_products(int M, int H, int lmax, int64_t n, int64_t nh, int nc,
                                                           const int *__restrict__ pj, const B *__restrict__ blocks,
                                                           const cd *__restrict__ alpha, const cd *__restrict__ x,
                                                           cd *__restrict__ part) {
    constexpr int NC = 4;
    const int p = blockIdx.x, lane = threadIdx.x;
    const int j = __builtin_amdgcn_readfirstlane(pj[p]);
    const B *blk = blocks + (int64_t)p * M * M;
    cd sum[RS][NC];
#pragma unroll
    for (int k = 0; k < RS; ++k)
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) sum[k][cc] = cd(0.0);
#pragma unroll 4
    for (int c = 0; c < M; ++c) {
        const int64_t xo = (int64_t)(c & 1) * nh + (int64_t)j * H + (c >> 1);
        cd xv[NC];
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) xv[cc] = cc < nc ? x[(int64_t)cc * n + xo] : cd(0.0);
#pragma unroll
        for (int k = 0; k < RS; ++k) {
            const cd b = block_value(blk[(int64_t)c * M + min(lane + 64 * k, M - 1)]);  // rows >= M unused
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) sum[k][cc] += b * xv[cc];
        }
    }
#pragma unroll
    for (int k = 0; k < RS; ++k) {
        const int r = lane + 64 * k;
        if (r < M) {
            const cd a = alpha[(int64_t)p * lmax + row_degree(r) - 1];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
                if (cc < nc) part[((int64_t)p * NC + cc) * M + r] = a * sum[k][cc];
        }
    }
}

// part[p][cc][r] as k_near_pair_products, from the same-helicity blocks [p][h][a][b]
// stored as bfloat16 pairs: one workgroup per pair, wave h takes the helicity-h rows,
// lane l the destinations b = l + 64 k (reference row 2 b + h); the sum runs over the H
// sources a of that helicity.  The source values x_j (both helicities, up to 4
// columns, zero past nc) are staged in LDS as FP32 first, so the loop is branch-free:
// the block values of 8 sources are loaded before any is used (8 loads in flight per
// lane), then multiplied with FP32 accumulation.  This preconditioner serves the
// inner solves of mixed precision: a rounded M changes iteration counts at most,
// never the answer, which the FP64 residual certifies.
template <int RS>
__global__ __launch_bounds__(128) void k_near_pair_products_helicity(int H, int lmax, int64_t n, int64_t nh, int nc,
                                                                     const int *__restrict__ pj,
                                                                     const uint32_t *__restrict__ blocks,
                                                                     const cd *__restrict__ alpha,
                                                                     const cd *__restrict__ x, cd *__restrict__ part) {
    constexpr int NC = 4, U = 8;
    extern __shared__ float2 xsf[];  // [h][a][NC]
    const int p = blockIdx.x, lane = threadIdx.x & 63, M = 2 * H;
    const int h = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j = pj[p];
    for (int t = threadIdx.x; t < 2 * H * NC; t += blockDim.x) {
        const int cc = t % NC, ha = t / NC, hh = ha / H, a = ha - hh * H;
        const cd v = cc < nc ? x[(int64_t)cc * n + (int64_t)hh * nh + (int64_t)j * H + a] : cd(0.0);
        xsf[t] = make_float2((float)v.re, (float)v.im);
    }
    __syncthreads();
    const uint32_t *blk = blocks + ((int64_t)p * 2 + h) * H * H;
    const float2 *xh = xsf + h * H * NC;
    int dst[RS];
#pragma unroll
    for (int k = 0; k < RS; ++k) dst[k] = min(lane + 64 * k, H - 1);  // destinations >= H unused
    float2 sum[RS][NC];
#pragma unroll
    for (int k = 0; k < RS; ++k)
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) sum[k][cc] = make_float2(0.f, 0.f);
    auto mac = [&](int a, const uint32_t (&bw)[RS]) {
        float br[RS], bi[RS];
#pragma unroll
        for (int k = 0; k < RS; ++k) {
            br[k] = __uint_as_float(bw[k] << 16);
            bi[k] = __uint_as_float(bw[k] & 0xffff0000u);
        }
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
            const float2 xv = xh[a * NC + cc];
#pragma unroll
            for (int k = 0; k < RS; ++k) {
                sum[k][cc].x = fmaf(br[k], xv.x, fmaf(-bi[k], xv.y, sum[k][cc].x));
                sum[k][cc].y = fmaf(br[k], xv.y, fmaf(bi[k], xv.x, sum[k][cc].y));
            }
        }
    };
    int a0 = 0;
    for (; a0 + U <= H; a0 += U) {
        uint32_t bv[U][RS];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < RS; ++k) bv[u][k] = blk[(a0 + u) * H + dst[k]];
#pragma unroll
        for (int u = 0; u < U; ++u) mac(a0 + u, bv[u]);
    }
    for (; a0 < H; ++a0) {
        uint32_t bv[RS];
#pragma unroll
        for (int k = 0; k < RS; ++k) bv[k] = blk[a0 * H + dst[k]];
        mac(a0, bv);
    }
#pragma unroll
    for (int k = 0; k < RS; ++k) {
        const int bd = lane + 64 * k;
        if (bd < H) {
            const int r = 2 * bd + h;
            const cd al = alpha[(int64_t)p * lmax + row_degree(r) - 1];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
                if (cc < nc) part[((int64_t)p * NC + cc) * M + r] = al * cd((double)sum[k][cc].x, (double)sum[k][cc].y);
        }
    }
}

__global__ __launch_bounds__(256) void k_near_finish(int M, int H, int lmax, int64_t n, int64_t nh, int nc,
                                                     const int *__restrict__ row_start, const cd *__restrict__ part,
                                                     const cd *__restrict__ mie, const cd *__restrict__ x,
                                                     cd *__restrict__ y) {
    constexpr int NC = 4;
    const int i = blockIdx.x;
    const int p0 = row_start[i], p1 = row_start[i + 1];
    for (int t = threadIdx.x; t < nc * (M / 2); t += blockDim.x) {
        const int cc = t / (M / 2), b = t - cc * (M / 2);
        cd s0(0.0), s1(0.0);  // rows 2 b (helicity 0) and 2 b + 1 (helicity 1)
        for (int p = p0; p < p1; ++p) {
            const cd *pp = part + ((int64_t)p * NC + cc) * M;
            s0 += pp[2 * b];
            s1 += pp[2 * b + 1];
        }
        const int l = row_degree(2 * b);
        const cd *t2 = mie + (size_t)(i * lmax + l - 1) * 4;
        for (int h = 0; h < 2; ++h) {
            const cd v = t2[h * 2 + 0] * s0 + t2[h * 2 + 1] * s1;
            const int64_t idx = (int64_t)cc * n + (int64_t)h * nh + (int64_t)i * H + b;
            y[idx] = x[idx] + v;
        }
    }
}

// y_j = x_j + sum_{pairs p with source j} C_p^H D(conj alpha_p) z_{i_p}, z = T^H x.
__global__ void k_near_apply_adjoint(int N, int M, int H, int lmax, int64_t n, int64_t nh, const int *col_start,
                                     const int *col_list, const int *pi, const cd *blocks, const cd *alpha,
                                     const cd *z, const cd *x, cd *y) {
    const int j = blockIdx.x, col = blockIdx.y;
    const int c = threadIdx.x;
    if (c >= M) return;
    const cd *zc = z + (int64_t)col * n;
    cd acc(0.0);
    for (int q = col_start[j]; q < col_start[j + 1]; ++q) {
        const int p = col_list[q];
        const int i = pi[p];
        const cd *blk = blocks + (int64_t)p * M * M + (int64_t)c * M;
        for (int r = 0; r < M; ++r) {
            const int l = row_degree(r);
            acc += conj(blk[r]) * (conj(alpha[(int64_t)p * lmax + l - 1]) *
                                   zc[(int64_t)(r & 1) * nh + (int64_t)i * H + (r >> 1)]);
        }
    }
    const int64_t idx = (int64_t)col * n + (int64_t)(c & 1) * nh + (int64_t)j * H + (c >> 1);
    y[idx] = x[idx] + acc;
}

// P_p[r, col] = sum_c C_p[r, c] V_{j_p}[c, col]   (grid np, one thread per row)
__global__ void k_pair_products(int64_t np, int M, int H, int64_t n, int64_t nh, int cols, const int *pj,
                                const cd *blocks, const cd *V, cd *P) {
    const int64_t p = blockIdx.x;
    const int r = threadIdx.x;
    if (r >= M) return;
    const int j = pj[p];
    const cd *blk = blocks + p * M * M;
    for (int cc = 0; cc < cols; ++cc) {
        const cd *vc = V + (int64_t)cc * n;
        cd s(0.0);
        for (int c = 0; c < M; ++c) s += blk[(int64_t)c * M + r] * vc[(int64_t)(c & 1) * nh + (int64_t)j * H + (c >> 1)];
        P[(p * cols + cc) * M + r] = s;
    }
}

// dL/d(out_p): dL/d(alpha) = 2 sum_{rows of degree l} Z_i conj(P_p) / |V|^2 as (re, im)
// pairs, times d alpha / d out = 0.5 (1 - tanh^2 out) per component.
__global__ void k_alpha_grad(int64_t np, int M, int H, int lmax, int64_t n, int64_t nh, int cols, const int *pi,
                             const cd *Z, const cd *P, double scale, const float *out, float *dout) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int i = pi[p];
    for (int t = 0; t < kOut; ++t) dout[p * kOut + t] = 0.f;
    for (int l = 1; l <= lmax && l <= kOut / 2; ++l) {
        cd g(0.0);
        for (int b = l * l - 1; b < l * l + 2 * l; ++b)
            for (int h = 0; h < 2; ++h) {
                const int r = 2 * b + h;
                for (int cc = 0; cc < cols; ++cc)
                    g += Z[(int64_t)cc * n + (int64_t)h * nh + (int64_t)i * H + b] * conj(P[(p * cols + cc) * M + r]);
            }
        const double tr = tanh((double)out[p * kOut + 2 * (l - 1)]), ti = tanh((double)out[p * kOut + 2 * (l - 1) + 1]);
        dout[p * kOut + 2 * (l - 1)] = (float)(2.0 * scale * g.re * 0.5 * (1.0 - tr * tr));
        dout[p * kOut + 2 * (l - 1) + 1] = (float)(2.0 * scale * g.im * 0.5 * (1.0 - ti * ti));
    }
}

static LearnedState *state(tg_cluster *c) {
    if (!c->learned) c->learned = new LearnedState();
    return c->learned;
}

// K nearest neighbours of every sphere, near blocks and features (device).
// The KT nearest other spheres of every sphere i, selected by (squared distance,
// index) as the host selection (std::pair order), then listed by ascending index: pair
// p = i KT + k is (pi, pj) = (i, out[p]), with row_start[i] = i KT and the
// displacement disp[p] = pos_i - pos_j.  One wave per sphere: lane l keeps the best KT
// of the candidates j = l, l + 64, ... in registers (each candidate carried down a
// fully unrolled insertion chain), then KT rounds of a wave-wide arg-min over the
// lanes' heads merge them.  The squared distances are rounded as on the host
// ((0 + t0^2) + t1^2) + t2^2, no fused multiply-adds, so equal distances tie the same
// way.
template <int KT>
__global__ __launch_bounds__(64) void k_nearest(int N, const double *__restrict__ pos, int *__restrict__ out,
                                                 int *__restrict__ pi, int *__restrict__ row_start,
                                                 double *__restrict__ disp) {
    const int i = blockIdx.x, lane = threadIdx.x;
    double bd[KT];
    int bj[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        bd[k] = __builtin_inf();
        bj[k] = 0x7fffffff;
    }
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    for (int j = lane; j < N; j += 64) {
        if (j == i) continue;
        const double t0 = __dsub_rn(xi, pos[3 * j]), t1 = __dsub_rn(yi, pos[3 * j + 1]),
                     t2 = __dsub_rn(zi, pos[3 * j + 2]);
        double d = __dadd_rn(__dadd_rn(__dadd_rn(0.0, __dmul_rn(t0, t0)), __dmul_rn(t1, t1)), __dmul_rn(t2, t2));
        if (d < bd[KT - 1] || (d == bd[KT - 1] && j < bj[KT - 1])) {
            int jj = j;
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                if (d < bd[k] || (d == bd[k] && jj < bj[k])) {
                    const double td = bd[k];
                    const int tj = bj[k];
                    bd[k] = d;
                    bj[k] = jj;
                    d = td;
                    jj = tj;
                }
            }
        }
    }
    int win[KT];  // the winners, the same in every lane
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        double d = bd[0];
        int j = bj[0];
        for (int off = 32; off > 0; off >>= 1) {
            const double od = __shfl_xor(d, off);
            const int oj = __shfl_xor(j, off);
            if (od < d || (od == d && oj < j)) {
                d = od;
                j = oj;
            }
        }
        if (bj[0] == j) {  // this lane held the winner: drop it from its list
#pragma unroll
            for (int m = 0; m + 1 < KT; ++m) {
                bd[m] = bd[m + 1];
                bj[m] = bj[m + 1];
            }
            bd[KT - 1] = __builtin_inf();
            bj[KT - 1] = 0x7fffffff;
        }
        win[k] = j;
    }
    // ascending index (distinct values: a fixed compare-exchange network)
#pragma unroll
    for (int a = 1; a < KT; ++a)
#pragma unroll
        for (int b = a; b > 0; --b)
            if (win[b] < win[b - 1]) {
                const int t = win[b];
                win[b] = win[b - 1];
                win[b - 1] = t;
            }
    if (lane < KT) {
        int j = win[0];
#pragma unroll
        for (int k = 1; k < KT; ++k)
            if (lane == k) j = win[k];
        const int64_t p = (int64_t)i * KT + lane;
        out[p] = j;
        if (pi) pi[p] = i;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (disp) disp[3 * p + a] = __dsub_rn(pos[3 * i + a], pos[3 * j + a]);
    }
    if (row_start && lane == 0) {
        row_start[i] = i * KT;
        if (i == N - 1) row_start[N] = N * KT;
    }
}

// column lists of the near pairs (the adjoint apply's), built on the host from pj the
// first time they are needed
static int ensure_pair_columns(tg_cluster *c) {
    LearnedState *s = c->learned;
    if (s->have_cols) return TG_OK;
    const int N = c->N;
    const int64_t np = s->npairs;
    std::vector<int> pj(np), col_start(N + 1, 0), col_list(np);
    TG_HIP(hipMemcpyAsync(pj.data(), s->pj, sizeof(int) * np, hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    for (int64_t p = 0; p < np; ++p) col_start[pj[p] + 1]++;
    for (int i = 0; i < N; ++i) col_start[i + 1] += col_start[i];
    std::vector<int> fill(col_start.begin(), col_start.end() - 1);
    for (int64_t p = 0; p < np; ++p) col_list[fill[pj[p]]++] = (int)p;
    TG_HIP(hipMemcpy(const_cast<int *>(s->col_start), col_start.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
    TG_HIP(hipMemcpy(const_cast<int *>(s->col_list), col_list.data(), sizeof(int) * np, hipMemcpyHostToDevice));
    s->have_cols = true;
    return TG_OK;
}

// features of the near pairs and room for the network's output; the blocks themselves
// are built on first use, in the form that use needs
static int pair_setup_tail(tg_cluster *c, LearnedState *s) {
    const int64_t np = s->npairs;
    s->have_blocks = false;
    s->have_blocksh = false;
    TG_TRY(s->feats.ensure(sizeof(float) * (size_t)np * kFeat));
    k_pair_features<<<(unsigned)((np + 127) / 128), 128, 0, c->stream>>>(
        np, s->pi, s->pj, c->b_pos.as<double>(), c->b_mie.as<cd>(), c->lmax, c->k0, s->feats.as<float>());
    TG_HIP(hipGetLastError());
    TG_TRY(s->alpha.ensure(sizeof(cd) * (size_t)np * c->lmax));
    TG_TRY(s->out16.ensure(sizeof(float) * (size_t)np * kOut));
    return TG_OK;
}

int learned_pairs(tg_cluster *c, int K) {
    TG_TRY(require_spheres(c, "the learned preconditioner"));
    LearnedState *s = state(c);
    if (s->npairs > 0 && s->K == K) return TG_OK;
    const int N = c->N;
    K = std::min(K, N - 1);
    std::vector<int> pi, pj;
    if (K == 8) {
        // on the device, pair lists, row starts and displacements included (the host
        // selection is O(N^2) and idles the GPU for ms); the column lists wait for the
        // first adjoint apply
        const int64_t np = (int64_t)N * 8;
        TG_TRY(s->meta.ensure(sizeof(int) * (size_t)(3 * np + 2 * (N + 1))));
        int *base = s->meta.as<int>();
        s->pi = base;
        s->pj = base + np;
        s->row_start = base + 2 * np;
        s->col_start = s->row_start + N + 1;
        s->col_list = s->col_start + N + 1;
        s->npairs = np;
        s->K = K;
        s->have_alpha = false;
        s->have_cols = false;
        TG_TRY(s->disp.ensure(sizeof(double) * 3 * (size_t)np));
        k_nearest<8><<<(unsigned)N, 64, 0, c->stream>>>(N, c->b_pos.as<double>(), base + np, base,
                                                        base + 2 * np, s->disp.as<double>());
        TG_HIP(hipGetLastError());
        return pair_setup_tail(c, s);
    } else {
        std::vector<std::pair<double, int>> dist(N);
        for (int i = 0; i < N; ++i) {
            int m = 0;
            for (int j = 0; j < N; ++j) {
                if (j == i) continue;
                double d = 0;
                for (int a = 0; a < 3; ++a) {
                    double t = c->positions[3 * i + a] - c->positions[3 * j + a];
                    d += t * t;
                }
                dist[m++] = {d, j};
            }
            std::partial_sort(dist.begin(), dist.begin() + K, dist.begin() + m);
            std::vector<int> nb;
            for (int t = 0; t < K; ++t) nb.push_back(dist[t].second);
            std::sort(nb.begin(), nb.end());
            for (int j : nb) {
                pi.push_back(i);
                pj.push_back(j);
            }
        }
    }
    const int64_t np = (int64_t)pi.size();
    std::vector<int> row_start(N + 1, 0), col_start(N + 1, 0), col_list(np);
    for (int64_t p = 0; p < np; ++p) {
        row_start[pi[p] + 1]++;
        col_start[pj[p] + 1]++;
    }
    for (int i = 0; i < N; ++i) {
        row_start[i + 1] += row_start[i];
        col_start[i + 1] += col_start[i];
    }
    std::vector<int> fill(col_start.begin(), col_start.end() - 1);
    for (int64_t p = 0; p < np; ++p) col_list[fill[pj[p]]++] = (int)p;
    std::vector<int> host;
    host.insert(host.end(), pi.begin(), pi.end());
    host.insert(host.end(), pj.begin(), pj.end());
    host.insert(host.end(), row_start.begin(), row_start.end());
    host.insert(host.end(), col_start.begin(), col_start.end());
    host.insert(host.end(), col_list.begin(), col_list.end());
    TG_TRY(s->meta.ensure(sizeof(int) * host.size()));
    TG_HIP(hipMemcpy(s->meta.ptr, host.data(), sizeof(int) * host.size(), hipMemcpyHostToDevice));
    int *base = s->meta.as<int>();
    s->pi = base;
    s->pj = base + np;
    s->row_start = base + 2 * np;
    s->col_start = s->row_start + N + 1;
    s->col_list = s->col_start + N + 1;
    s->npairs = np;
    s->K = K;
    s->have_alpha = false;
    s->have_cols = true;
    // near blocks C_ij = translation at pos_i - pos_j
    std::vector<double> disp(3 * np);
    for (int64_t p = 0; p < np; ++p)
        for (int a = 0; a < 3; ++a) disp[3 * p + a] = c->positions[3 * pi[p] + a] - c->positions[3 * pj[p] + a];
    TG_TRY(s->disp.ensure(sizeof(double) * disp.size()));
    TG_HIP(hipMemcpy(s->disp.ptr, disp.data(), sizeof(double) * disp.size(), hipMemcpyHostToDevice));
    return pair_setup_tail(c, s);
}

int learned_set_output(tg_cluster *c, const float *out16_dev) {
    LearnedState *s = state(c);
    if (s->npairs == 0) {
        set_error("near pairs not built: call tg_cluster_near_pairs first");
        return TG_INVALID_INPUT;
    }
    if (out16_dev != s->out16.as<float>())
        TG_HIP(hipMemcpyAsync(s->out16.ptr, out16_dev, sizeof(float) * s->npairs * kOut, hipMemcpyDeviceToDevice,
                              c->stream));
    const int64_t cnt = s->npairs * c->lmax;
    k_alpha_from_output<<<(unsigned)((cnt + 255) / 256), 256, 0, c->stream>>>(s->npairs, c->lmax, s->out16.as<float>(),
                                                                               s->alpha.as<cd>());
    TG_HIP(hipGetLastError());
    s->have_alpha = true;
    return TG_OK;
}

// c128 reference-layout blocks (adjoint, training, FP64 solves)
static int ensure_blocks128(tg_cluster *c) {
    LearnedState *s = c->learned;
    if (s->have_blocks) return TG_OK;
    const int M = c->M;
    TG_TRY(s->blocks.ensure(sizeof(cd) * (size_t)s->npairs * M * M));
    TG_TRY(launch_translation_blocks(c->lmax, c->plan, c->k0, s->npairs, s->disp.as<double>(), 0,
                                     s->blocks.as<cd>(), c->stream));
    s->have_blocks = true;
    return TG_OK;
}

template <class B>
static int near_apply_forward(tg_cluster *c, const B *blocks, const cd *x, cd *y, int cols) {
    LearnedState *s = c->learned;
    const int M = c->M;
    TG_TRY(s->part.ensure(sizeof(cd) * (size_t)s->npairs * 4 * M));
    cd *part = s->part.as<cd>();
    for (int c0 = 0; c0 < cols; c0 += 4) {
        const int nc = std::min(4, cols - c0);
        const cd *xc = x + (int64_t)c0 * c->n;
#define TG_NEAR(RS)                                                                                              \
    k_near_pair_products<RS, B><<<(unsigned)s->npairs, 64, 0, c->stream>>>(M, c->H, c->lmax, c->n, c->nh, nc,   \
                                                                             s->pj, blocks, s->alpha.as<cd>(), xc, \
                                                                             part)
        const int rs = (M + 63) / 64;
        if (rs <= 1) TG_NEAR(1);
        else if (rs == 2) TG_NEAR(2);
        else if (rs <= 4) TG_NEAR(4);
        else TG_NEAR(6);
#undef TG_NEAR
        TG_HIP(hipGetLastError());
        k_near_finish<<<c->N, 256, 0, c->stream>>>(M, c->H, c->lmax, c->n, c->nh, nc, s->row_start, part,
                                                   c->b_mie.as<cd>(), xc, y + (int64_t)c0 * c->n);
        TG_HIP(hipGetLastError());
    }
    return TG_OK;
}

// forward apply with the same-helicity near blocks as bfloat16 pairs (an eighth of the
// c128 blocks' bytes; FP32 products and sums), for the inner solves of mixed precision
static int learned_near_apply_inner(tg_cluster *c, const cd *x, cd *y, int cols) {
    LearnedState *s = c->learned;
    const int H = c->H, M = c->M;
    if (!s->have_blocksh) {
        TG_TRY(s->blocksh.ensure(sizeof(uint32_t) * (size_t)s->npairs * 2 * H * H));
        TG_TRY(launch_translation_blocks_helicity_bf16(c->lmax, c->plan, c->k0, s->npairs, s->disp.as<double>(),
                                                       s->blocksh.as<uint32_t>(), c->stream,
                                                       fp32_tables_in_range(c)));
        s->have_blocksh = true;
    }
    TG_TRY(s->part.ensure(sizeof(cd) * (size_t)s->npairs * 4 * M));
    cd *part = s->part.as<cd>();
    for (int c0 = 0; c0 < cols; c0 += 4) {
        const int nc = std::min(4, cols - c0);
        const cd *xc = x + (int64_t)c0 * c->n;
        const int rs = (H + 63) / 64;
#define TG_NEARH(RS)                                                                                               \
    k_near_pair_products_helicity<RS><<<(unsigned)s->npairs, 128, sizeof(float2) * 2 * H * 4, c->stream>>>(           \
        H, c->lmax, c->n, c->nh, nc, s->pj, s->blocksh.as<uint32_t>(), s->alpha.as<cd>(), xc, part)
        if (rs <= 1) TG_NEARH(1);
        else if (rs == 2) TG_NEARH(2);
        else if (rs <= 4) TG_NEARH(4);
        else TG_NEARH(6);
#undef TG_NEARH
        TG_HIP(hipGetLastError());
        k_near_finish<<<c->N, 256, 0, c->stream>>>(M, H, c->lmax, c->n, c->nh, nc, s->row_start, part,
                                                   c->b_mie.as<cd>(), xc, y + (int64_t)c0 * c->n);
        TG_HIP(hipGetLastError());
    }
    return TG_OK;
}

int learned_near_apply(tg_cluster *c, const cd *x, cd *y, int cols, int adjoint) {
    LearnedState *s = c->learned;
    const int M = c->M;
    TG_TRY(ensure_blocks128(c));
    if (!adjoint) return near_apply_forward(c, s->blocks.as<cd>(), x, y, cols);
    TG_TRY(ensure_pair_columns(c));
    TG_TRY(s->work.ensure(sizeof(cd) * (size_t)cols * c->n));
    cd *z = s->work.as<cd>();
    TG_TRY(local_apply(c, x, z, cols, 1, nullptr));
    k_near_apply_adjoint<<<dim3(c->N, cols), ((M + 63) / 64) * 64, 0, c->stream>>>(
        c->N, M, c->H, c->lmax, c->n, c->nh, s->col_start, s->col_list, s->pi, s->blocks.as<cd>(), s->alpha.as<cd>(),
        z, x, y);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

int learned_setup(tg_cluster *c, Preconditioner *) {
    LearnedState *s = state(c);
    if (c->mlp) {
        TG_TRY(learned_pairs(c, s->K > 0 ? s->K : 8));
        TG_TRY(mlp_forward(c->mlp, s->feats.as<float>(), s->npairs, s->out16.as<float>()));
        TG_TRY(learned_set_output(c, s->out16.as<float>()));
        return TG_OK;
    }
    if (s->npairs > 0 && s->have_alpha) return TG_OK;
    set_error("learned preconditioner: attach a network (tg_cluster_set_mlp) or set the pair weights");
    return TG_INVALID_INPUT;
}

int learned_apply(tg_cluster *c, Preconditioner *, const cd *x, cd *y, int cols, int adjoint, int storage) {
    if (!adjoint && storage == TG_STORE_C64) return learned_near_apply_inner(c, x, y, cols);
    return learned_near_apply(c, x, y, cols, adjoint);
}

void learned_destroy(Preconditioner *) {}

// near pairs, features, weights and blocks to be rebuilt, their buffers kept (a fresh
// block from the buffer cache synchronises the device)
void learned_invalidate(tg_cluster *c) {
    if (!c->learned) return;
    LearnedState *s = c->learned;
    s->npairs = 0;
    s->have_alpha = false;
    s->have_blocks = false;
    s->have_blocksh = false;
    s->have_cols = false;
}

void learned_release(tg_cluster *c) {
    if (!c->learned) return;
    LearnedState *s = c->learned;
    for (DevBuf *b : {&s->meta, &s->disp, &s->blocks, &s->blocksh, &s->part, &s->feats, &s->out16, &s->alpha,
                      &s->work})
        b->release();
    delete s;
    c->learned = nullptr;
}

}  // namespace tg

using namespace tg;

extern "C" {

int tg_cluster_near_pairs(tg_cluster *c, int k_nearest, int64_t *npairs) {
    if (!c || k_nearest < 1) {
        set_error("need a cluster and k_nearest >= 1");
        return TG_INVALID_INPUT;
    }
    if (c->N < 2) {
        set_error("a single sphere has no near pairs");
        return TG_INVALID_INPUT;
    }
    TG_TRY(learned_pairs(c, k_nearest));
    if (npairs) *npairs = c->learned->npairs;
    return TG_OK;
}

int tg_cluster_pair_list(tg_cluster *c, int32_t *pairs_host) {
    if (!c || !c->learned || c->learned->npairs == 0) {
        set_error("near pairs not built");
        return TG_INVALID_INPUT;
    }
    const int64_t np = c->learned->npairs;
    std::vector<int> a(np), b(np);
    TG_HIP(hipMemcpy(a.data(), c->learned->pi, sizeof(int) * np, hipMemcpyDeviceToHost));
    TG_HIP(hipMemcpy(b.data(), c->learned->pj, sizeof(int) * np, hipMemcpyDeviceToHost));
    for (int64_t p = 0; p < np; ++p) {
        pairs_host[2 * p] = a[p];
        pairs_host[2 * p + 1] = b[p];
    }
    return TG_OK;
}

int tg_cluster_pair_features(tg_cluster *c, float *features_dev) {
    if (!c || !c->learned || c->learned->npairs == 0) {
        set_error("near pairs not built");
        return TG_INVALID_INPUT;
    }
    TG_HIP(hipMemcpyAsync(features_dev, c->learned->feats.ptr, sizeof(float) * c->learned->npairs * kFeat,
                          hipMemcpyDeviceToDevice, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    return TG_OK;
}

int tg_cluster_set_pair_output(tg_cluster *c, const float *out_dev) {
    if (!c) {
        set_error("null cluster");
        return TG_INVALID_INPUT;
    }
    TG_TRY(learned_set_output(c, out_dev));
    if (c->pc_kind == TG_PC_LEARNED) destroy_preconditioner(c);
    TG_HIP(hipStreamSynchronize(c->stream));
    return TG_OK;
}

int tg_cluster_set_mlp(tg_cluster *c, tg_mlp *m) {
    if (!c) {
        set_error("null cluster");
        return TG_INVALID_INPUT;
    }
    c->mlp = m;
    if (c->pc_kind == TG_PC_LEARNED) destroy_preconditioner(c);
    return TG_OK;
}

int tg_cluster_learned_apply(tg_cluster *c, const tg_c128 *x, tg_c128 *y, int cols, int adjoint) {
    if (!c || !c->learned || !c->learned->have_alpha || cols < 1) {
        set_error("pair weights not set");
        return TG_INVALID_INPUT;
    }
    const int64_t count = (int64_t)cols * c->n;
    DevBuf io;
    TG_TRY(io.ensure(sizeof(cd) * count * 2));
    cd *xs = io.as<cd>(), *ys = xs + count;
    int st = ref_to_split(c, (const cd *)x, xs, cols);
    if (!st) st = adjoint == 2 ? learned_near_apply_inner(c, xs, ys, cols) : learned_near_apply(c, xs, ys, cols, adjoint);
    if (!st) st = split_to_ref(c, ys, (cd *)y, cols);
    if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = TG_DEVICE_ERROR;
    io.release();
    return st;
}

int tg_cluster_learned_loss(tg_cluster *c, const tg_c128 *probes, int cols, double *loss, float *dout_dev) {
    if (!c || !c->learned || !c->learned->have_alpha || cols < 1 || !loss) {
        set_error("pair weights not set");
        return TG_INVALID_INPUT;
    }
    LearnedState *s = c->learned;
    const int64_t count = (int64_t)cols * c->n;
    const int M = c->M;
    DevBuf io;
    TG_TRY(io.ensure(sizeof(cd) * (count * 4 + (size_t)s->npairs * cols * M)));
    cd *V = io.as<cd>(), *U = V + count, *R = U + count, *Z = R + count, *P = Z + count;
    int st = TG_OK;
    do {
        if ((st = ref_to_split(c, (const cd *)probes, V, cols))) break;
        if ((st = learned_near_apply(c, V, U, cols, 0))) break;          // U = M^{-1} V
        if ((st = apply_operator(c, U, R, cols, 0, TG_OP_MATRIX_FREE_F64))) break;
        if ((st = vec_axpby(c->stream, cd(-1.0), V, cd(1.0), R, count))) break;  // R = A U - V
        std::vector<double> vn(cols), rn(cols);
        if ((st = vec_norms(c, V, c->n, cols, c->n, vn.data()))) break;
        if ((st = vec_norms(c, R, c->n, cols, c->n, rn.data()))) break;
        double v2 = 0, r2 = 0;
        for (int k = 0; k < cols; ++k) {
            v2 += vn[k] * vn[k];
            r2 += rn[k] * rn[k];
        }
        *loss = r2 / v2;
        if (!dout_dev) break;
        if ((st = apply_operator(c, R, U, cols, 1, TG_OP_MATRIX_FREE_F64))) break;  // A^H R
        if ((st = local_apply(c, U, Z, cols, 1, nullptr))) break;                     // Z = T^H A^H R
        if ((st = ensure_blocks128(c))) break;
        k_pair_products<<<(unsigned)s->npairs, ((M + 63) / 64) * 64, 0, c->stream>>>(
            s->npairs, M, c->H, c->n, c->nh, cols, s->pj, s->blocks.as<cd>(), V, P);
        if (hipGetLastError() != hipSuccess) { st = TG_DEVICE_ERROR; break; }
        k_alpha_grad<<<(unsigned)((s->npairs + 127) / 128), 128, 0, c->stream>>>(
            s->npairs, M, c->H, c->lmax, c->n, c->nh, cols, s->pi, Z, P, 1.0 / v2, s->out16.as<float>(), dout_dev);
        if (hipGetLastError() != hipSuccess) { st = TG_DEVICE_ERROR; break; }
    } while (0);
    if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = TG_DEVICE_ERROR;
    if (st == TG_DEVICE_ERROR && tg_last_error()[0] == 0) set_error("learned loss kernels failed");
    io.release();
    return st;
}

}  // extern "C"
