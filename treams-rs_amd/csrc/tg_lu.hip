// Dense LU of the cluster operator A = I - T C on the device, in c64 (refined in FP64)
// or c128: replaces Lu::new / solve_in_place / solve_adjoint_in_place
// (linalg/mod.rs:398-531) behind InteractionFactor (cluster/interaction.rs:180-240)
// and sphere_cluster_factor (cluster/spheres.rs:85-94).
//
// Storage: A in the helicity-split order (rows and columns h N H + i H + b),
// column-major, leading dimension n, assembled directly by the pair kernel.
// Blocked right-looking LU with block width NB = 64: the diagonal block is factored
// in LDS by one wave with partial pivoting inside the block, its row swaps are
// applied to the whole block row, the panels below and to the right are solved
// against U11 / L11, and the trailing matrix is updated by a complex GEMM in the 3M
// form (three real products per complex product) on MFMA:
// v_mfma_f32_16x16x4f32 for c64, v_mfma_f64_16x16x4f64 for c128, FP32 / FP64
// accumulation.  Pivoting inside diagonal blocks keeps every update a GEMM; the
// reference pivots over the whole column (faer).  The operator's identity self
// blocks make the diagonal blocks well conditioned for separated spheres, and every
// solve is certified by the FP64 residual b - A x: an unstable factorisation shows
// up as a refinement that fails with TG_NOT_CONVERGED, never as a wrong answer.
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <vector>

#include "tg_gmres.h"
#include "tg_internal.h"
#include "tg_tangent.h"

#include "tg_lu.h"
#include "tg_mfma.h"

struct tg_factor {
    tg_cluster *cluster = nullptr;
    int precision = TG_STORE_C128;
    int64_t n = 0;
    tg::DevBuf lu;   // the factors, n x n column-major
    tg::LuSet set;   // a batch of one matrix
    // equilibration: the factors are those of D_r A D_c; eq holds D_r[n], D_c[n] (FP64
    // powers of two, split order)
    bool equilibrated = false;
    tg::DevBuf eq;
    tg::DevBuf eqwork;  // scratch of the equilibration statistics
    // c128 factors of at most kDenseResidual unknowns keep the FP64 operator as
    // assembled (before equilibration): the refinement residual of forward solves is
    // then one GEMM instead of a matrix-free apply
    tg::DevBuf op;
    // pinned host squares of the right-hand side and residual norms of a solve: the
    // right-hand side's are read with the first residual's (one synchronisation)
    tg::HostBuf norms;
    // pivoting: 0 inside 64-row diagonal blocks, 1 over whole columns (GETRF order)
    int pivoting = 0;
    // fallbacks taken: 1 when a stalled or singular solve refactored the operator (in
    // c128 with whole-column pivoting)
    int fallback = 0;
    // a factor of a caller's dense matrix (no cluster): `cluster` is a bare handle
    // carrying the stream, owned by the factor; `op` always holds the matrix
    bool dense = false;
};

namespace tg {

constexpr int NB = 64;
constexpr int64_t kDenseResidual = 4096;  // largest operator copied for the residual (256 MiB)

// ------------------------------------------------------------ factor kernels --
//
// Every kernel works on a batch of matrices (grid.z = matrix z): matrix z has
// dimension bn[z], starts at element boff[z] of the base pointer, and its rows (the
// pivots and the packed vectors of the solves) start at bvoff[z].  A single matrix is
// a batch of one.  Workgroups beyond a matrix's own dimension return at once.

struct Batch {
    const int64_t *off;   // element offset of each matrix
    const int *n;         // dimension of each matrix
    const int64_t *voff;  // row offset of each matrix in the packed vectors and pivots
};

// Value of x in lane `lane` (wave-uniform) for every lane.
__device__ __forceinline__ float read_lane(float x, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane));
}
__device__ __forceinline__ double read_lane(double x, int lane) {
    const long long bits = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)bits, lane), hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <class T>
__device__ __forceinline__ cx<T> read_lane(cx<T> x, int lane) {
    return cx<T>(read_lane(x.re, lane), read_lane(x.im, lane));
}

// 1 / z with one real division: z scaled exactly by the power of two 2^-e that brings
// its larger component into [0.5, 1), so |z 2^-e|^2 lies in [0.25, 2] whatever |z|.
template <class T>
__device__ __forceinline__ cx<T> recip_scaled(cx<T> z) {
    int e;
    (void)frexp(fmax(fabs(z.re), fabs(z.im)), &e);
    const T a = ldexp(z.re, -e), b = ldexp(z.im, -e);
    const T r = (T)1 / (a * a + b * b);
    return cx<T>(ldexp(a * r, -e), ldexp(-b * r, -e));
}

// v -= l r as four FMAs.
template <class T>
__device__ __forceinline__ void sub_mul(cx<T> &v, cx<T> l, cx<T> r) {
    v.re = fma(-l.im, -r.im, fma(-l.re, r.re, v.re));
    v.im = fma(-l.im, r.re, fma(-l.re, r.im, v.im));
}

// Largest value over the wave (wave-uniform): DPP steps inside each row of 16 lanes,
// then the four row maxima by readlane.  Every lane of the wave must be active.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    auto mx = [](uint32_t a, uint32_t b) { return a > b ? a : b; };
    v = mx(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    v = mx(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    v = mx(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = mx(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false));  // row_mirror
    const uint32_t a = __builtin_amdgcn_readlane(v, 15), b = __builtin_amdgcn_readlane(v, 31);
    const uint32_t c = __builtin_amdgcn_readlane(v, 47), d = __builtin_amdgcn_readlane(v, 63);
    return mx(mx(a, b), mx(c, d));
}

// Factor the nb x nb diagonal block at k0 with partial pivoting inside the block;
// piv[k0 + k] = row exchanged with row k0 + k.  W waves: lane r holds row r, wave w
// the columns w, w + W, w + 2W, ... in registers (dealt cyclically, so that every
// wave has about the same number of columns right of any step, and two waves share a
// SIMD).  Rows never move during the factorisation (implicit pivoting): ord = the step
// at which a row became the pivot (-1 before), and each row lands at position ord
// when the block is stored, which is the arrangement of the equivalent exchange
// sequence piv (tracked in cur = the row at each position).  The pivot of column k is
// found by the wave owning it (one 32-bit wave max over keys = the high word of |a|^2
// with the low 6 bits replaced by 63 - row: magnitudes within 2^-14 of the largest
// tie, and the lowest row wins), which publishes it with the multipliers in LDS
// (double buffered: one barrier per step).  Look-ahead: after the barrier of step k
// the owner of column k + 1 applies step k to that column first and finds its pivot
// before its other updates (branch-free, so the two interleave); every wave updates
// its columns right of k (multipliers zero on the rows already pivoted), reading the
// pivot row with readlane.  Each entry sees the same operations in the same order
// whatever W.  Measured at lmax 3 (one 64 x 64 block): 2.8 us of loads, stores and
// launch, 13.6 us of the 64 barrier / LDS steps alone, 9 us of pivot searches and
// 12 us of updates; a rolled step loop (slots picked by uniform index) is slower.
template <class T>
__global__ __launch_bounds__(512) void k_lu_diag(cx<T> *A0, Batch bt, int64_t k0, int *piv0, int *status) {
    constexpr int W = 8, J = NB / W;
    const int z = blockIdx.z;
    const int64_t n = bt.n[z];
    if (k0 >= n) return;
    cx<T> *A = A0 + bt.off[z];
    int *piv = piv0 + bt.voff[z];
    const int nb = (int)min((int64_t)NB, n - k0);
    // wave-uniform (a scalar register): the column tests below are scalar branches
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    __shared__ lcx<T> lmul[2][NB];
    __shared__ int spiv[2];
    cx<T> v[J];
    const int64_t row = k0 + min(lane, nb - 1);
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int c = W * j + wave;
        const cx<T> e = A[(k0 + min(c, nb - 1)) * n + row];
        v[j] = (lane < nb && c < nb) ? e : cx<T>();
    }
    int ord = lane < nb ? -1 : NB;  // rows beyond the block never pivot
    int cur = lane;
    // pivot of column k (slot q of the calling wave) among the rows not yet pivoted;
    // a zero pivot publishes -1 and zero multipliers
    auto pivot = [&](int k, int q) {
        const int buf = k & 1;
        const double m = (double)abs2(v[q]);
        const uint32_t key =
            ord < 0 ? ((uint32_t)((uint64_t)__double_as_longlong(m) >> 32) & ~63u) | (uint32_t)(63 - lane) : 0u;
        const int p = 63 - (int)(wave_max_u32(key) & 63u);
        const cx<T> pv = read_lane(v[q], p);
        const bool ok = (double)abs2(pv) > 0.0;  // uniform over the wave
        const cx<T> inv = recip_scaled(ok ? pv : cx<T>((T)1));
        const bool below = ok && ord < 0 && lane != p;
        const cx<T> l = below ? v[q] * inv : cx<T>();
        lmul[buf][lane].set(l);
        if (below) v[q] = l;  // the stored L entry
        if (lane == 0) spiv[buf] = ok ? p : -1;
    };
    if (wave == 0) pivot(0, 0);
    // fully unrolled (the register slots are indexed by step): steps past nb are skipped
#pragma clang loop unroll(full)
    for (int k = 0; k < NB; ++k) {
        if (k >= nb) continue;
        const int buf = k & 1, owner = k % W, q = k / W;
        const int k1 = k + 1, owner1 = k1 % W, q1 = k1 / W;
        __syncthreads();
        // both LDS reads issued together (the multipliers are unused after a zero pivot)
        const int praw = spiv[buf];
        const cx<T> l = lmul[buf][lane].get();
        const int p = __builtin_amdgcn_readfirstlane(praw);
        if (p < 0) {
            if (threadIdx.x == 0) status[z] = 1;
            return;
        }
        if (lane == p) ord = k;
        // the exchange sequence: position j of row p swaps with position k
        const int j = __ffsll((long long)__ballot(cur == p)) - 1;
        const int ck = __builtin_amdgcn_readlane(cur, k);
        cur = lane == j ? ck : (lane == k ? p : cur);
        if (threadIdx.x == 0) piv[k0 + k] = (int)(k0 + j);
        // l = 0 on the rows already pivoted and on the pivot row: they keep their
        // values (v - 0 rp = v for finite pivot rows).  Columns right of k: slots from q
        // on in the waves after the owner, from q + 1 on in the owner and the waves
        // before it; in the owner of column k + 1 (a wave after the owner, or wave 0
        // when k + 1 starts a new round) they start at that column's slot q1.
        if (wave == owner1 && k1 < nb) {
            sub_mul(v[q1], l, read_lane(v[q1], p));
            pivot(k1, q1);
#pragma unroll
            for (int jj = q1 + 1; jj < J; ++jj) sub_mul(v[jj], l, read_lane(v[jj], p));
        } else if (wave > owner) {
#pragma unroll
            for (int jj = q; jj < J; ++jj) sub_mul(v[jj], l, read_lane(v[jj], p));
        } else {
#pragma unroll
            for (int jj = q + 1; jj < J; ++jj) sub_mul(v[jj], l, read_lane(v[jj], p));
        }
    }
    if (lane < nb) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int c = W * j + wave;
            if (c < nb) A[(k0 + c) * n + k0 + ord] = v[j];
        }
    }
}

// Local source row of row k0 + lane after the block's exchanges (row k0 + k with
// piv[k0 + k], clamped to the block, for k = 0 .. nb - 1, in order), composed into one
// gather: lane i of the wave holds the composition's entry i, each exchange swaps two
// lanes' entries.
__device__ inline int block_perm(const int *piv, int64_t k0, int nb, int lane) {
    const int pv = lane < nb ? (int)(min(max((int64_t)piv[k0 + lane], k0), k0 + nb - 1) - k0) : lane;
    int v = lane;
    for (int k = 0; k < nb; ++k) {
        const int p = __builtin_amdgcn_readlane(pv, k);
        const int a = __builtin_amdgcn_readlane(v, k), b = __builtin_amdgcn_readlane(v, p);
        v = lane == k ? b : (lane == p ? a : v);
    }
    return v;
}

// Apply the row exchanges of the block at k0 to the columns [cb, ce) outside the
// block: one wave per column, lane i gathering row k0 + perm(i) of the column (one
// load for the wave, then one store: the column is read whole before it is written).
// Pivots are clamped to the block's rows: after a zero pivot the rest of the block's
// entries are never written, and the factorisation reports TG_SINGULAR.
template <class T>
__global__ __launch_bounds__(256) void k_lu_swap(cx<T> *A0, Batch bt, int64_t k0, const int *piv0, int64_t cb,
                                                 int64_t ce) {
    const int z = blockIdx.z;
    const int64_t n = bt.n[z];
    if (k0 >= n) return;
    const int nb = (int)min((int64_t)NB, n - k0);
    const int lane = threadIdx.x & 63;
    const int src = block_perm(piv0 + bt.voff[z], k0, nb, lane);
    const int64_t cend = min(ce, n), waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    cx<T> *A = A0 + bt.off[z];
    for (int64_t col = cb + (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); col < cend;
         col += waves) {
        if (col >= k0 && col < k0 + nb) continue;  // uniform over the wave
        cx<T> *c = A + col * n + k0;
        if (src != lane) {
            const cx<T> x = c[src];
            c[lane] = x;
        }
    }
}

// The deferred exchanges of the L columns: block k0 (from the second block pair on;
// blockIdx.y selects it) exchanges rows of the columns of every earlier block pair.
// Blocks cover disjoint rows, so their exchanges commute and run concurrently.
template <class T>
__global__ __launch_bounds__(256) void k_lu_swap_left(cx<T> *A0, Batch bt, const int *piv0) {
    const int z = blockIdx.z;
    const int64_t n = bt.n[z];
    const int64_t k0 = (int64_t)(2 + blockIdx.y) * NB;
    if (k0 >= n) return;
    const int nb = (int)min((int64_t)NB, n - k0);
    const int lane = threadIdx.x & 63;
    const int src = block_perm(piv0 + bt.voff[z], k0, nb, lane);
    const int64_t cend = (k0 / (2 * NB)) * (2 * NB), waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    cx<T> *A = A0 + bt.off[z];
    for (int64_t col = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); col < cend; col += waves) {
        cx<T> *c = A + col * n + k0;
        if (src != lane) {
            const cx<T> x = c[src];
            c[lane] = x;
        }
    }
}

// L21 = A21 U11^{-1}, in place.  Each wave solves R consecutive rows below the block
// with lane c holding column c of every row (for c128 one 128-byte segment of each
// column); step k scales entry k by 1 / u_kk and broadcasts it (readlane), and the
// columns right of k subtract x_k u_kc, with U11 staged in LDS by rows (lane c reads
// u_kc).  A chain of 64 short steps per wave, all rows in parallel; the sums run over
// k in ascending order.  R = 8 amortises the staging of U11 over many rows; with few
// rows below the block, R = 2 spreads them over four times the workgroups and
// shortens each step of the chain.
template <class T, int R>
__global__ __launch_bounds__(256) void k_trsm_lower_panel(cx<T> *A0, Batch bt, int64_t k0) {
    const int z = blockIdx.z;
    const int64_t n = bt.n[z];
    const int64_t rw = k0 + NB + (int64_t)blockIdx.x * 4 * R;  // first row of the workgroup
    if (k0 + NB >= n || rw >= n) return;                         // uniform over the workgroup
    cx<T> *A = A0 + bt.off[z];
    __shared__ lcx<T> U[NB][NB];  // U[k][c] = u_kc
    for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
        const int col = t / NB, row = t - col * NB;
        U[row][col].set(A[(k0 + col) * n + k0 + row]);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t r0 = rw + wave * R;
    const cx<T> rd = cx<T>((T)1) / U[lane][lane].get();
    cx<T> a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = A[(k0 + lane) * n + min(r0 + r, n - 1)];
#pragma unroll 8
    for (int k = 0; k < NB; ++k) {
        const cx<T> u = U[k][lane].get();
        const cx<T> rk = read_lane(rd, k);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const cx<T> xk = read_lane(a[r], k) * rk;
            if (lane == k)
                a[r] = xk;
            else if (lane > k)
                a[r] -= xk * u;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (r0 + r < n) A[(k0 + lane) * n + r0 + r] = a[r];
}

// U12 = L11^{-1} A12 (unit lower L11 of the block at k0) for the columns [cb, ce), in
// place: one wave per column, lane r holding row r; each step broadcasts the solved
// entry and updates the rows below it.
template <class T>
__global__ __launch_bounds__(256) void k_trsm_upper_panel(cx<T> *A0, Batch bt, int64_t k0, int64_t cb, int64_t ce) {
    const int z = blockIdx.z;
    const int64_t n = bt.n[z];
    if (k0 + NB >= n || cb >= n) return;
    const int64_t cend = min(ce, n);
    cx<T> *A = A0 + bt.off[z];
    const int nb = NB;
    __shared__ lcx<T> L[NB][NB];  // [col][row]
    for (int t = threadIdx.x; t < nb * nb; t += blockDim.x) {
        const int col = t / nb, row = t - col * nb;
        L[col][row].set(A[(k0 + col) * n + k0 + row]);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t col = cb + (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); col < cend;
         col += waves) {
        cx<T> *c = A + col * n + k0;
        cx<T> v = c[lane];
#pragma unroll 8
        for (int k = 0; k < nb - 1; ++k) {
            const cx<T> ck = read_lane(v, k);
            if (lane > k) v -= L[k][lane].get() * ck;
        }
        c[lane] = v;
    }
}

// A[rb : re, cb : ce] -= A[rb : re, k0 : k0 + K] A[k0 : k0 + K, cb : ce] (rows and
// columns clipped to n; K a multiple of 16 with k0 + K <= n wherever a tile is left)
// with the 3M complex product on MFMA, one 64 x 64 tile per workgroup; rows and
// columns beyond the range are never written.
template <class T>
__global__ __launch_bounds__(256, 2) void k_lu_update(cx<T> *A0, Batch bt, int64_t k0, int K, int64_t rb, int64_t re,
                                                      int64_t cb, int64_t ce) {
    const int z = blockIdx.z;
    const int64_t nz = bt.n[z];
    const int64_t rend = min(re, nz), cend = min(ce, nz);
    const int64_t r0 = rb + (int64_t)blockIdx.y * 64, c0 = cb + (int64_t)blockIdx.x * 64;
    if (r0 >= rend || c0 >= cend) return;  // uniform over the workgroup
    const int64_t n = nz;
    cx<T> *A = A0 + bt.off[z];
    __shared__ TileLds<T> sm;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32, rr = lane & 15;
    // c64: the tile's A22 values are loaded before the K loop (their latency overlaps the
    // MFMAs); c128 loads them afterwards, which keeps two waves per SIMD without spills
    constexpr bool prefetch = sizeof(T) == 4;
    cx<T> creg[2][2][4];
    auto load_c = [&] {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int64_t gr = min(r0 + wm + a * 16 + rr, rend - 1);
                    const int64_t gc = min(c0 + wn + b * 16 + mfma_out_row<T>(lane, v), cend - 1);
                    creg[a][b][v] = A[gc * n + gr];
                }
    };
    if (prefetch) load_c();
    Tile3M<T> acc;
    // L21 = A[:, k0 : k0 + K] (rows absolute), U12 = A[k0 : k0 + K, :] (columns absolute)
    zgemm_tile_3m<T, false>(A + k0 * n, n, A + k0, n, r0, c0, rend, cend, K, sm, acc);
    if (!prefetch) load_c();
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t gr = r0 + wm + a * 16 + rr, gc = c0 + wn + b * 16 + mfma_out_row<T>(lane, v);
                const T t1 = acc.t1[a][b][v], t2 = acc.t2[a][b][v], t3 = acc.t3[a][b][v];
                cx<T> d = creg[a][b][v];
                d.re -= t1 - t2;
                d.im -= t3 - t1 - t2;
                if (gr < rend && gc < cend) A[gc * n + gr] = d;
            }
}

// ------------------------------------------------------------- solve kernels --
//
// Vectors are packed: column q of the batch is x + q ld, and matrix z's rows start at
// x + q ld + voff[z].

// Row permutation of `cols` packed vector columns: dst[i] = src[perm[i]] (forward,
// P x) or dst[perm[i]] = src[i] (inverse, P^T x); perm holds packed indices.
__global__ void k_lu_permute(const cd *src, cd *dst, int64_t ld, int cols, const int *perm, int inverse) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ld * cols) return;
    const int64_t col = t / ld, i = t - col * ld;
    const int64_t p = perm[i];
    if (!inverse)
        dst[col * ld + i] = src[col * ld + p];
    else
        dst[col * ld + p] = src[col * ld + i];
}

// Triangular solve with the diagonal block at k0: 8-column workgroups (grid.x), wave
// w solving vector column 8 blockIdx.x + w, lane r holding row r (FP64 vectors,
// factors in T).  WHICH = 0: L y = b (unit lower); 1: U x = y; 2: U^H z = g;
// 3: L^H w = z (unit upper).  The block is staged in LDS oriented so that step k
// reads S[k][lane] (conflict-free): S[k][r] = A(r, k) for WHICH 0 / 1 and
// S[k][r] = conj(A(k, r)) for WHICH 2 / 3.  The substitution is unrolled by 8, so
// the LDS reads of 8 steps are issued ahead of their readlane / FMA chain.
template <class T, int WHICH>
__global__ __launch_bounds__(512) void k_lu_block_solve(const cx<T> *A0, Batch bt, int64_t k0, cd *x0, int64_t ld,
                                                        int cols) {
    const int z = blockIdx.z;
    const int64_t n = bt.n[z];
    if (k0 >= n) return;
    const cx<T> *A = A0 + bt.off[z];
    const int nb = (int)min((int64_t)NB, n - k0);
    __shared__ lcx<T> S[NB][NB];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr bool trans = WHICH >= 2;
#pragma unroll 8
    for (int t = threadIdx.x; t < nb * nb; t += 512) {
        const int c = t / nb, r = t - c * nb;  // coalesced over r
        const cx<T> e = A[(k0 + c) * n + k0 + r];
        if (!trans)
            S[c][r].set(e);
        else
            S[r][c].set(cx<T>(e.re, -e.im));  // conj(A(r, c)) at S[r][c]: (A^H)(c, r)
    }
    __syncthreads();
    const int c0 = blockIdx.x * 8;
    if (wave >= cols - c0) return;
    cd *v = x0 + (int64_t)(c0 + wave) * ld + bt.voff[z] + k0;
    cd val = lane < nb ? v[lane] : cd();
    auto s = [&](int k) {
        const cx<T> e = S[k][lane].get();
        return cd((double)e.re, (double)e.im);
    };
    // the reciprocal of this lane's diagonal entry, for the non-unit triangles (every
    // lane at once, off the substitution's latency chain)
    cd rdiag(0.0);
    if ((WHICH == 1 || WHICH == 2) && lane < nb) {
        const cx<T> e = S[lane][lane].get();
        rdiag = cd(1.0) / cd((double)e.re, (double)e.im);
    }
    if (WHICH == 0 || WHICH == 2) {
        // lower triangular (L, or U^H with S[k][r] = conj(U(k, r)))
#pragma unroll 8
        for (int k = 0; k < nb; ++k) {
            if (WHICH == 2 && lane == k) val = val * rdiag;
            const cd vk = read_lane(val, k);
            if (lane > k && lane < nb) val -= s(k) * vk;
        }
    } else {
        // upper triangular (U, or L^H with S[k][r] = conj(L(k, r)))
#pragma unroll 8
        for (int k = nb - 1; k >= 0; --k) {
            if (WHICH == 1 && lane == k) val = val * rdiag;
            const cd vk = read_lane(val, k);
            if (lane < k) val -= s(k) * vk;
        }
    }
    if (lane < nb) v[lane] = val;
}

// Off-diagonal part of the solves for up to 8 vector columns:
// which = 0: x[r] -= sum_k A[r, k0 + k] x[k0 + k] for r >= k0 + nb (L21);
// which = 1: x[r] -= sum_k A[r, k0 + k] x[k0 + k] for r < k0 (U01);
// one thread per row, 16 loads in flight, coalesced along the panel's columns.
template <class T>
__global__ __launch_bounds__(256) void k_lu_panel_rows(const cx<T> *A0, Batch bt, int64_t k0, cd *x0, int64_t ld,
                                                       int cols, int which) {
    const int z = blockIdx.z;
    const int64_t n = bt.n[z];
    if (k0 >= n) return;
    const int nb = (int)min((int64_t)NB, n - k0);
    const int64_t rcount = which == 0 ? n - k0 - nb : k0;
    const int64_t rr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ((int64_t)blockIdx.x * blockDim.x >= rcount) return;  // uniform over the workgroup
    const cx<T> *A = A0 + bt.off[z];
    cd *x = x0 + bt.voff[z];
    __shared__ lcx<double> seg[8][NB];
    // entries beyond nb are zero, so the loads below need no bounds branch
    for (int t = threadIdx.x; t < cols * NB; t += blockDim.x) {
        const int q = t / NB, k = t % NB;
        seg[q][k].set(k < nb ? x[(int64_t)q * ld + k0 + k] : cd(0.0));
    }
    __syncthreads();
    if (rr >= rcount) return;
    const int64_t r = which == 0 ? k0 + nb + rr : rr;
    cd s[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] = cd(0.0);
    for (int k = 0; k < nb; k += 16) {
        cx<T> e[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) e[u] = A[(k0 + min(k + u, nb - 1)) * n + r];  // clamped: seg is 0 there
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const cd ev((double)e[u].re, (double)e[u].im);
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (q < cols) s[q] += ev * seg[q][k + u].get();
        }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (q < cols) x[(int64_t)q * ld + r] -= s[q];
}

// which = 2: x[c] -= sum_k conj(A[k0 + k, c]) x[k0 + k] for c >= k0 + nb (U12^H);
// which = 3: x[c] -= sum_k conj(A[k0 + k, c]) x[k0 + k] for c < k0 (L10^H);
// one wave per column c: lane k reads A[k0 + k, c] (a contiguous segment) and the
// wave reduces the products with shuffles.
template <class T>
__global__ __launch_bounds__(256) void k_lu_panel_cols(const cx<T> *A0, Batch bt, int64_t k0, cd *x0, int64_t ld,
                                                       int cols, int which) {
    const int z = blockIdx.z;
    const int64_t n = bt.n[z];
    if (k0 >= n) return;
    const int nb = (int)min((int64_t)NB, n - k0);
    const int64_t ccount = which == 2 ? n - k0 - nb : k0;
    if ((int64_t)blockIdx.x * (blockDim.x >> 6) >= ccount) return;  // uniform over the workgroup
    const cx<T> *A = A0 + bt.off[z];
    cd *x = x0 + bt.voff[z];
    __shared__ lcx<double> seg[8][NB];
    for (int t = threadIdx.x; t < cols * nb; t += blockDim.x)
        seg[t / nb][t % nb].set(x[(int64_t)(t / nb) * ld + k0 + t % nb]);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t cc = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); cc < ccount; cc += waves) {
        const int64_t c = which == 2 ? k0 + nb + cc : cc;
        cx<T> e = lane < nb ? A[c * n + k0 + lane] : cx<T>();
        const cd ev((double)e.re, -(double)e.im);
        cd s[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) s[q] = q < cols ? ev * seg[q][lane < nb ? lane : 0].get() : cd(0.0);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                s[q].re += __shfl_xor(s[q].re, off);
                s[q].im += __shfl_xor(s[q].im, off);
            }
        if (lane < cols) {
            cd sq = s[0];
#pragma unroll
            for (int q = 1; q < 8; ++q)
                if (q == lane) sq = s[q];
            x[(int64_t)lane * ld + c] -= sq;
        }
    }
}

// ------------------------------------------------------------------ host side --

int luset_init(LuSet &ls, int precision, void *A, const std::vector<int64_t> &off, const std::vector<int> &n) {
    ls.precision = precision;
    ls.A = A;
    ls.count = (int)n.size();
    ls.nmax = 0;
    ls.ntot = 0;
    ls.n_h = n;
    ls.voff_h.assign(ls.count, 0);
    for (int z = 0; z < ls.count; ++z) {
        ls.voff_h[z] = ls.ntot;
        ls.ntot += n[z];
        ls.nmax = std::max(ls.nmax, n[z]);
    }
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_off = 0, o_n = al(sizeof(int64_t) * ls.count), o_voff = al(o_n + sizeof(int) * ls.count);
    const size_t o_piv = al(o_voff + sizeof(int64_t) * ls.count);
    const size_t o_perm = al(o_piv + sizeof(int) * ls.ntot);
    const size_t o_stat = al(o_perm + sizeof(int) * ls.ntot);
    const size_t total = al(o_stat + sizeof(int) * ls.count);
    TG_TRY(ls.meta.ensure(total));
    std::vector<char> host(o_piv, 0);
    memcpy(host.data() + o_off, off.data(), sizeof(int64_t) * ls.count);
    memcpy(host.data() + o_n, n.data(), sizeof(int) * ls.count);
    memcpy(host.data() + o_voff, ls.voff_h.data(), sizeof(int64_t) * ls.count);
    char *base = ls.meta.as<char>();
    TG_HIP(hipMemcpy(base, host.data(), o_piv, hipMemcpyHostToDevice));
    ls.off = (const int64_t *)(base + o_off);
    ls.n = (const int *)(base + o_n);
    ls.voff = (const int64_t *)(base + o_voff);
    ls.piv = (int *)(base + o_piv);
    ls.perm = (int *)(base + o_perm);
    ls.status = (int *)(base + o_stat);
    TG_TRY(ls.work.ensure(sizeof(cd) * 8 * (size_t)ls.ntot));
    return TG_OK;
}

// Look-ahead streams and their event pairs, kept per device for the next factor (a
// stream's creation costs about as much as a small operator's whole factorisation).
namespace {
struct PanelSet {
    hipStream_t stream;
    hipEvent_t ready, done;
};
std::mutex g_panel_mu;
std::map<int, std::vector<PanelSet>> g_panels;
}  // namespace

static int panel_acquire(LuSet &ls) {
    if (ls.panel) return TG_OK;
    int dev = 0;
    TG_HIP(hipGetDevice(&dev));
    ls.panel_dev = dev;
    {
        std::lock_guard<std::mutex> lock(g_panel_mu);
        auto &v = g_panels[dev];
        if (!v.empty()) {
            ls.panel = v.back().stream;
            ls.ready = v.back().ready;
            ls.done = v.back().done;
            v.pop_back();
            return TG_OK;
        }
    }
    TG_HIP(hipStreamCreateWithFlags(&ls.panel, hipStreamNonBlocking));
    TG_HIP(hipEventCreateWithFlags(&ls.ready, hipEventDisableTiming));
    TG_HIP(hipEventCreateWithFlags(&ls.done, hipEventDisableTiming));
    return TG_OK;
}

void luset_release(LuSet &ls) {
    if (ls.panel) {
        const bool ok = hipStreamSynchronize(ls.panel) == hipSuccess && ls.ready && ls.done;
        if (ok) {
            std::lock_guard<std::mutex> lock(g_panel_mu);
            g_panels[ls.panel_dev].push_back(PanelSet{ls.panel, ls.ready, ls.done});
        } else {
            (void)hipStreamDestroy(ls.panel);
            if (ls.ready) (void)hipEventDestroy(ls.ready);
            if (ls.done) (void)hipEventDestroy(ls.done);
        }
        ls.panel = nullptr;
        ls.ready = ls.done = nullptr;
    }
    ls.meta.release();
    ls.work.release();
    ls.many.release();
}

// The packed row permutation of the solves, perm[voff + k0 + i] = voff + k0 + (the
// block's composed exchanges)(i): pivots stay inside their 64-row block, so each block
// composes its own; one wave per block (grid.x), matrix grid.z.
__global__ __launch_bounds__(64) void k_lu_perm(Batch bt, const int *piv0, int *perm0) {
    const int z = blockIdx.z;
    const int64_t n = bt.n[z], k0 = (int64_t)blockIdx.x * NB;
    if (k0 >= n) return;
    const int nb = (int)min((int64_t)NB, n - k0);
    const int lane = threadIdx.x;
    const int src = block_perm(piv0 + bt.voff[z], k0, nb, lane);
    if (lane < nb) perm0[bt.voff[z] + k0 + lane] = (int)(bt.voff[z] + k0 + src);
}

template <class T>
static int luset_factor_t(LuSet &ls, hipStream_t s) {
    cx<T> *A = reinterpret_cast<cx<T> *>(ls.A);
    const Batch bt{ls.off, ls.n, ls.voff};
    const unsigned z = (unsigned)ls.count;
    const int64_t nmax = ls.nmax;
    TG_HIP(hipMemsetAsync(ls.status, 0, sizeof(int) * ls.count, s));
    // Block pairs a = [p0, p1), b = [p1, p2): the trailing matrix right of b is updated
    // once with K = 128 instead of twice with K = 64 (halving its read-modify-write
    // traffic, the update's HBM bound).  Look-ahead: the panel of pair j + 1 (its two
    // diagonal blocks, their panel solves and the update between them) runs on a second
    // stream while the main stream updates the columns right of that panel.  Row
    // exchanges reach the panel's own columns at once, the columns right of the pair
    // on the main stream before their panel solves, and the L columns left of the
    // pair in one pass at the end.
    int st = TG_OK;
    auto check = [&](hipError_t e) {
        if (e != hipSuccess && st == TG_OK) {
            set_error("HIP error %s in the LU factorisation", hipGetErrorString(e));
            st = TG_DEVICE_ERROR;
        }
    };
    // small operators: everything on the main stream (the look-ahead's event waits cost
    // more there than the overlap saves)
    const bool ahead = nmax > 1024;
    if (ahead && !ls.panel && (st = panel_acquire(ls))) return st;
    hipStream_t ps = ahead ? ls.panel : s;
    hipEvent_t ready = ls.ready, done = ls.done;
    auto join = [&](hipEvent_t e, hipStream_t from, hipStream_t to) {  // to waits for from
        if (!ahead) return;
        check(hipEventRecord(e, from));
        check(hipStreamWaitEvent(to, e, 0));
    };
    auto swap = [&](hipStream_t q, int64_t k, int64_t cb, int64_t ce) {  // one wave per column
        if (ce > cb)
            k_lu_swap<T><<<dim3((unsigned)std::min<int64_t>((ce - cb + 3) / 4, 2048), 1, z), 256, 0, q>>>(
                A, bt, k, ls.piv, cb, ce);
    };
    auto diag = [&](hipStream_t q, int64_t k) {
        k_lu_diag<T><<<dim3(1, 1, z), 512, 0, q>>>(A, bt, k, ls.piv, ls.status);
    };
    auto lower = [&](hipStream_t q, int64_t k) {  // rows below the block at k
        const int64_t rows = nmax - k - NB;
        if (rows <= 4096)
            k_trsm_lower_panel<T, 2><<<dim3((unsigned)((rows + 7) / 8), 1, z), 256, 0, q>>>(A, bt, k);
        else
            k_trsm_lower_panel<T, 8><<<dim3((unsigned)((rows + 31) / 32), 1, z), 256, 0, q>>>(A, bt, k);
    };
    auto upper = [&](hipStream_t q, int64_t k, int64_t cb, int64_t ce) {
        k_trsm_upper_panel<T><<<dim3((unsigned)std::min<int64_t>((ce - cb + 3) / 4, 1024), 1, z), 256, 0, q>>>(
            A, bt, k, cb, ce);
    };
    auto update = [&](hipStream_t q, int64_t k, int K, int64_t rb, int64_t re, int64_t cb, int64_t ce) {
        if (ce > cb && re > rb)
            k_lu_update<T><<<dim3((unsigned)((ce - cb + 63) / 64), (unsigned)((re - rb + 63) / 64), z), 256, 0, q>>>(
                A, bt, k, K, rb, re, cb, ce);
    };
    if (st == TG_OK) join(ready, s, ps);  // the panel stream starts after the work before
    for (int64_t p0 = 0; p0 < nmax && st == TG_OK; p0 += 2 * NB) {
        const int64_t p1 = p0 + NB, p2 = p1 + NB, p4 = p2 + 2 * NB;
        const int64_t pe = std::min(p2, nmax);
        // the panel [p0, p2), on the panel stream
        diag(ps, p0);
        swap(ps, p0, p0, pe);
        if (p1 < nmax) {
            lower(ps, p0);
            upper(ps, p0, p1, p2);                  // U_ab
            update(ps, p0, NB, p1, nmax, p1, p2);   // block column b below block a
            diag(ps, p1);
            swap(ps, p1, p0, pe);                   // includes L_ba
            if (p2 < nmax) lower(ps, p1);
        }
        check(hipGetLastError());
        join(done, ps, s);
        if (p2 >= nmax) continue;
        // right of the panel, on the main stream
        swap(s, p0, p2, nmax);
        swap(s, p1, p2, nmax);
        upper(s, p0, p2, nmax);                     // U12 of block a
        update(s, p0, NB, p1, p2, p2, nmax);        // rows of b: A12_b -= L_ba U12_a
        upper(s, p1, p2, nmax);                     // U12 of block b
        update(s, p0, 2 * NB, p2, nmax, p2, std::min(p4, nmax));  // the next panel first
        join(ready, s, ps);
        update(s, p0, 2 * NB, p2, nmax, p4, nmax);  // overlaps the next panel
        check(hipGetLastError());
    }
    if (st == TG_OK && nmax > 2 * NB) {
        // blocks from the second pair on (y), a wave per column of the earlier pairs (x)
        const unsigned blocks = (unsigned)((nmax + NB - 1) / NB - 2);
        k_lu_swap_left<T><<<dim3((unsigned)std::min<int64_t>((nmax + 3) / 4, 64), blocks, z), 256, 0, s>>>(A, bt,
                                                                                                         ls.piv);
    }
    check(hipGetLastError());
    if (st == TG_OK) {
        // the panel stream's last work joins the main stream (the last pair's panel
        // has no main-stream work after it)
        join(done, ps, s);
    } else if (ahead) {
        (void)hipStreamSynchronize(ps);
    }
    if (st != TG_OK) return st;
    // each matrix's exchanges composed into one gather over the packed rows:
    // (P x)[voff + i] = x[voff + perm_local[i]]
    k_lu_perm<<<dim3((unsigned)((nmax + NB - 1) / NB), 1, z), 64, 0, s>>>(bt, ls.piv, ls.perm);
    TG_HIP(hipGetLastError());
    std::vector<int> stat(ls.count);
    TG_HIP(hipMemcpyAsync(stat.data(), ls.status, sizeof(int) * ls.count, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    for (int v : stat)
        if (v) {
            set_error("zero pivot in the LU factorisation");
            return TG_SINGULAR;
        }
    return TG_OK;
}

int luset_factor(LuSet &ls, hipStream_t s) {
    return ls.precision == TG_STORE_C64 ? luset_factor_t<float>(ls, s) : luset_factor_t<double>(ls, s);
}

// ------------------------------------------ whole-column pivoting (GETRF order) --
// The factorisation of one matrix with partial pivoting over whole columns
// (linalg/mod.rs:398-428 through faer's partial-pivot LU): a 64-column panel is
// factored column by column, each pivot searched over all remaining rows, the rows
// exchanged across the panel; then the panel's exchanges reach the other columns, the
// block row right of the panel is solved with the panel's unit lower triangle, and
// the trailing matrix is updated by the MFMA tile of the main LU (k_lu_update).  The
// factors land in the same storage and the exchanges in the same gather permutation
// as the main LU's, so every solve kernel serves both.  It is the fallback when
// pivoting inside 64-row diagonal blocks meets a zero pivot or its solves stall: three
// launches per column, so it is a latency-bound path, not the fast one.

// Pivot of column j: the row of largest |a_ij|^2 over i in [j, n), the lowest row on
// ties; one workgroup.  A zero (or NaN) pivot sets *status.
template <class T>
__global__ __launch_bounds__(1024) void k_gepp_pivot(const cx<T> *A, int64_t n, int64_t j, int *pivot, int *status) {
    __shared__ double sv[1024];
    __shared__ int si[1024];
    const int tid = threadIdx.x;
    double best = -1.0;
    int bi = (int)j;
    const cx<T> *col = A + j * n;
    for (int64_t i = j + tid; i < n; i += 1024) {
        const double re = (double)col[i].re, im = (double)col[i].im;
        const double m = re * re + im * im;
        if (m > best) {  // rows ascend per thread: the first of equal maxima stays
            best = m;
            bi = (int)i;
        }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (tid < w) {
            const double a = sv[tid], b = sv[tid + w];
            if (b > a || (b == a && si[tid + w] < si[tid])) {
                sv[tid] = b;
                si[tid] = si[tid + w];
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        pivot[j] = si[0];
        if (!(sv[0] > 0.0)) *status = 1;
    }
}

// Exchange rows j and pivot[j] across the panel columns [k0, k1) (k1 - k0 <= 64).
template <class T>
__global__ __launch_bounds__(64) void k_gepp_swap_panel(cx<T> *A, int64_t n, int64_t j, int64_t k0, int64_t k1,
                                                        const int *pivot) {
    const int64_t c = k0 + threadIdx.x, p = pivot[j];
    if (c >= k1 || p == j) return;
    cx<T> *col = A + c * n;
    const cx<T> t = col[j];
    col[j] = col[p];
    col[p] = t;
}

// Column j of the panel: l_i = a_ij / a_jj for i > j, then a_ic -= l_i a_jc for the
// panel columns j < c < k1; one thread per row.
template <class T>
__global__ __launch_bounds__(256) void k_gepp_eliminate(cx<T> *A, int64_t n, int64_t j, int64_t k1) {
    __shared__ T ur[64][2];
    __shared__ T rpv[2];
    const int nc = (int)(k1 - j - 1);
    for (int q = threadIdx.x; q < nc; q += blockDim.x) {
        const cx<T> u = A[(j + 1 + q) * n + j];
        ur[q][0] = u.re;
        ur[q][1] = u.im;
    }
    if (threadIdx.x == 0) {
        const cx<T> r = recip_scaled(A[j * n + j]);
        rpv[0] = r.re;
        rpv[1] = r.im;
    }
    __syncthreads();
    const cx<T> rp(rpv[0], rpv[1]);
    const int64_t i = j + 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const cx<T> l = A[j * n + i] * rp;
    A[j * n + i] = l;
    for (int q = 0; q < nc; ++q) sub_mul(A[(j + 1 + q) * n + i], l, cx<T>(ur[q][0], ur[q][1]));
}

// The panel's exchanges (pivot[k0 .. k1)) in order on every column outside the panel,
// [0, k0) and [k1, n); one thread per column.
template <class T>
__global__ void k_gepp_laswp(cx<T> *A, int64_t n, int64_t k0, int64_t k1, const int *pivot) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n - (k1 - k0)) return;
    const int64_t c = t < k0 ? t : t + (k1 - k0);
    cx<T> *col = A + c * n;
    for (int64_t j = k0; j < k1; ++j) {
        const int64_t p = pivot[j];
        if (p != j) {
            const cx<T> v = col[j];
            col[j] = col[p];
            col[p] = v;
        }
    }
}

// U12 = L11^{-1} A12 for the rows [k0, k1) of the columns [k1, n), L11 unit lower;
// forward substitution, one thread per column (L11 read through the caches).
template <class T>
__global__ __launch_bounds__(64) void k_gepp_trsm(cx<T> *A, int64_t n, int64_t k0, int64_t k1) {
    const int64_t c = k1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    cx<T> *col = A + c * n;
    const int kb = (int)(k1 - k0);
    for (int r = 1; r < kb; ++r) {
        cx<T> v = col[k0 + r];
        for (int q = 0; q < r; ++q) sub_mul(v, A[(k0 + q) * n + k0 + r], col[k0 + q]);
        col[k0 + r] = v;
    }
}

// perm[i] = the original row at position i after the exchanges pivot[0 .. n), in order.
__global__ void k_gepp_perm(const int *pivot, int64_t n, int *perm) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int64_t i = 0; i < n; ++i) perm[i] = (int)i;
    for (int64_t j = 0; j < n; ++j) {
        const int p = pivot[j];
        const int t = perm[j];
        perm[j] = perm[p];
        perm[p] = t;
    }
}

template <class T>
static int luset_factor_gepp_t(LuSet &ls, hipStream_t s) {
    if (ls.count != 1) {
        set_error("whole-column pivoting factors one matrix at a time");
        return TG_INVALID_INPUT;
    }
    cx<T> *A = reinterpret_cast<cx<T> *>(ls.A);
    const int64_t n = ls.n_h[0];
    DevBuf pv;
    TG_TRY(pv.ensure(sizeof(int) * (size_t)n));
    int *pivot = pv.as<int>();
    TG_HIP(hipMemsetAsync(ls.status, 0, sizeof(int), s));
    const Batch bt{ls.off, ls.n, ls.voff};
    auto blocks = [](int64_t w, int b) { return (unsigned)std::max<int64_t>(1, (w + b - 1) / b); };
    for (int64_t k0 = 0; k0 < n; k0 += NB) {
        const int64_t k1 = std::min<int64_t>(k0 + NB, n);
        for (int64_t j = k0; j < k1; ++j) {
            k_gepp_pivot<T><<<1, 1024, 0, s>>>(A, n, j, pivot, ls.status);
            k_gepp_swap_panel<T><<<1, 64, 0, s>>>(A, n, j, k0, k1, pivot);
            if (j + 1 < n) k_gepp_eliminate<T><<<blocks(n - j - 1, 256), 256, 0, s>>>(A, n, j, k1);
        }
        if (n > k1 - k0) k_gepp_laswp<T><<<blocks(n - (k1 - k0), 256), 256, 0, s>>>(A, n, k0, k1, pivot);
        if (k1 < n) {
            k_gepp_trsm<T><<<blocks(n - k1, 64), 64, 0, s>>>(A, n, k0, k1);
            k_lu_update<T><<<dim3(blocks(n - k1, 64), blocks(n - k1, 64), 1), 256, 0, s>>>(A, bt, k0, (int)NB, k1, n,
                                                                                       k1, n);
        }
        TG_HIP(hipGetLastError());
    }
    k_gepp_perm<<<1, 64, 0, s>>>(pivot, n, ls.perm);
    TG_HIP(hipGetLastError());
    int stat = 0;
    TG_HIP(hipMemcpyAsync(&stat, ls.status, sizeof(int), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    pv.release();
    if (stat) {
        set_error("zero pivot in the LU factorisation with whole-column pivoting (singular operator)");
        return TG_SINGULAR;
    }
    return TG_OK;
}

int luset_factor_gepp(LuSet &ls, hipStream_t s) {
    return ls.precision == TG_STORE_C64 ? luset_factor_gepp_t<float>(ls, s) : luset_factor_gepp_t<double>(ls, s);
}

// x (packed rows, FP64, `cols` <= 8 columns of stride ld = ntot) <- A^{-1} x or A^{-H} x.
template <class T>
static int luset_solve8(LuSet &ls, hipStream_t s, cd *x, int cols, int adjoint) {
    const cx<T> *A = reinterpret_cast<const cx<T> *>(ls.A);
    const Batch bt{ls.off, ls.n, ls.voff};
    const unsigned z = (unsigned)ls.count;
    const int64_t nmax = ls.nmax, ld = ls.ntot;
    cd *work = ls.work.as<cd>();
    auto grid = [](int64_t w) { return (unsigned)std::max<int64_t>(1, (w + 255) / 256); };
    // one wave per column, at most 2048 workgroups of 4 waves (grid-stride beyond)
    auto wgrid = [](int64_t w) { return (unsigned)std::min<int64_t>(2048, std::max<int64_t>(1, (w + 3) / 4)); };
    const int64_t last = ((nmax - 1) / NB) * NB;
    if (!adjoint) {
        // P A = L U: x <- P x, L y = x, U x = y
        k_lu_permute<<<grid(ld * cols), 256, 0, s>>>(x, work, ld, cols, ls.perm, 0);
        TG_HIP(hipMemcpyAsync(x, work, sizeof(cd) * ld * cols, hipMemcpyDeviceToDevice, s));
        for (int64_t k0 = 0; k0 < nmax; k0 += NB) {
            k_lu_block_solve<T, 0><<<dim3(1, 1, z), 512, 0, s>>>(A, bt, k0, x, ld, cols);
            if (nmax - k0 - NB > 0)
                k_lu_panel_rows<T><<<dim3(grid(nmax - k0 - NB), 1, z), 256, 0, s>>>(A, bt, k0, x, ld, cols, 0);
        }
        for (int64_t k0 = last; k0 >= 0; k0 -= NB) {
            k_lu_block_solve<T, 1><<<dim3(1, 1, z), 512, 0, s>>>(A, bt, k0, x, ld, cols);
            if (k0 > 0) k_lu_panel_rows<T><<<dim3(grid(k0), 1, z), 256, 0, s>>>(A, bt, k0, x, ld, cols, 1);
        }
    } else {
        // A^H = U^H L^H P: U^H z = g, L^H w = z, x = P^T w
        for (int64_t k0 = 0; k0 < nmax; k0 += NB) {
            k_lu_block_solve<T, 2><<<dim3(1, 1, z), 512, 0, s>>>(A, bt, k0, x, ld, cols);
            if (nmax - k0 - NB > 0)
                k_lu_panel_cols<T><<<dim3(wgrid(nmax - k0 - NB), 1, z), 256, 0, s>>>(A, bt, k0, x, ld, cols, 2);
        }
        for (int64_t k0 = last; k0 >= 0; k0 -= NB) {
            k_lu_block_solve<T, 3><<<dim3(1, 1, z), 512, 0, s>>>(A, bt, k0, x, ld, cols);
            if (k0 > 0) k_lu_panel_cols<T><<<dim3(wgrid(k0), 1, z), 256, 0, s>>>(A, bt, k0, x, ld, cols, 3);
        }
        k_lu_permute<<<grid(ld * cols), 256, 0, s>>>(x, work, ld, cols, ls.perm, 1);
        TG_HIP(hipMemcpyAsync(x, work, sizeof(cd) * ld * cols, hipMemcpyDeviceToDevice, s));
    }
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// X[rb : re, :] -= F[rb : re, k0 : k0 + K] X[k0 : k0 + K, :] for `cols` packed FP64
// columns (matrix z's rows at voff[z]), F the factors in T (rows below the block: L;
// rows above: U), on FP64 MFMA in the 3M form, one 64 x 64 tile of X per workgroup.
template <class T>
__global__ __launch_bounds__(256, 2) void k_lu_solve_update(const cx<T> *A0, Batch bt, int64_t k0, int64_t rb_,
                                                            int64_t re_, cd *x0, int64_t ldx, int cols) {
    const int z = blockIdx.z;
    const int64_t n = bt.n[z];
    if (k0 >= n) return;
    const int K = (int)min((int64_t)NB, n - k0);
    const int64_t re = min(re_, n), r0 = rb_ + (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
    if (r0 >= re || c0 >= cols) return;
    const cx<T> *F = A0 + bt.off[z];
    cd *X = x0 + bt.voff[z];
    __shared__ TileLds<double> sm;
    Tile3M<double> acc;
    const int64_t cmax = cols;
    auto bcol = [=](int64_t c) { return (const cd *)(X + min(c, cmax - 1) * ldx + k0); };
    zgemm_tile_3m_cols<double, true, 64, 64>(F + k0 * n, n, bcol, r0, c0, re, K, sm, acc);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32, rr = lane & 15;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t gr = r0 + wm + a * 16 + rr, gc = c0 + wn + b * 16 + mfma_out_row<double>(lane, v);
                if (gr >= re || gc >= cols) continue;
                const double t1 = acc.t1[a][b][v], t2 = acc.t2[a][b][v], t3 = acc.t3[a][b][v];
                cd &y = X[gc * ldx + gr];
                y = cd(y.re - (t1 - t2), y.im - (t3 - t1 - t2));
            }
}

// Forward solve of many columns: x <- P x, L y = x, U x = y, each 64-row block solved
// by 8-column workgroups and the rest of the rows updated by one MFMA GEMM per block.
template <class T>
static int luset_solve_many(LuSet &ls, hipStream_t s, cd *x, int cols) {
    const cx<T> *A = reinterpret_cast<const cx<T> *>(ls.A);
    const Batch bt{ls.off, ls.n, ls.voff};
    const unsigned z = (unsigned)ls.count;
    const int64_t nmax = ls.nmax, ld = ls.ntot;
    // kept with the set (a block released here would make the next solve synchronise)
    DevBuf &tmp = ls.many;
    TG_TRY(tmp.ensure(sizeof(cd) * (size_t)ld * cols));
    const unsigned pgrid = (unsigned)((ld * cols + 255) / 256);
    k_lu_permute<<<pgrid, 256, 0, s>>>(x, tmp.as<cd>(), ld, cols, ls.perm, 0);
    TG_HIP(hipMemcpyAsync(x, tmp.as<cd>(), sizeof(cd) * ld * cols, hipMemcpyDeviceToDevice, s));
    const unsigned cchunks = (unsigned)((cols + 7) / 8), ctiles = (unsigned)((cols + 63) / 64);
    for (int64_t k0 = 0; k0 < nmax; k0 += NB) {
        k_lu_block_solve<T, 0><<<dim3(cchunks, 1, z), 512, 0, s>>>(A, bt, k0, x, ld, cols);
        const int64_t below = nmax - k0 - NB;
        if (below > 0)
            k_lu_solve_update<T><<<dim3(ctiles, (unsigned)((below + 63) / 64), z), 256, 0, s>>>(
                A, bt, k0, k0 + NB, nmax, x, ld, cols);
    }
    const int64_t last = ((nmax - 1) / NB) * NB;
    for (int64_t k0 = last; k0 >= 0; k0 -= NB) {
        k_lu_block_solve<T, 1><<<dim3(cchunks, 1, z), 512, 0, s>>>(A, bt, k0, x, ld, cols);
        if (k0 > 0)
            k_lu_solve_update<T><<<dim3(ctiles, (unsigned)((k0 + 63) / 64), z), 256, 0, s>>>(A, bt, k0, 0, k0, x, ld,
                                                                                            cols);
    }
    TG_HIP(hipGetLastError());
    return TG_OK;
}

int luset_solve(LuSet &ls, hipStream_t s, cd *x, int cols, int adjoint) {
    if (!adjoint && cols > 8)
        return ls.precision == TG_STORE_C64 ? luset_solve_many<float>(ls, s, x, cols)
                                            : luset_solve_many<double>(ls, s, x, cols);
    for (int c0 = 0; c0 < cols; c0 += 8) {
        const int w = std::min(8, cols - c0);
        cd *xc = x + (int64_t)c0 * ls.ntot;
        TG_TRY(ls.precision == TG_STORE_C64 ? luset_solve8<float>(ls, s, xc, w, adjoint)
                                            : luset_solve8<double>(ls, s, xc, w, adjoint));
    }
    return TG_OK;
}

// ------------------------------------------------------------ equilibration --
//
// linalg/mod.rs:137-236 (equilibrate): an operator whose entries span more than a
// factor 10 while some diagonal entry is below a tenth of the largest gets its rows
// and columns scaled by the reciprocals of their largest magnitudes (max(|re|, |im|)),
// the axis with the smaller minimum first, the other axis measured after it.  Here the
// scales are rounded to powers of two, so scaling is exact in either precision; the
// reference's further power-of-two sum balancing (linalg/mod.rs:238-300) is not done.

__device__ __forceinline__ double eq_mag(cd z) { return fmax(fabs(z.re), fabs(z.im)); }
__device__ __forceinline__ double eq_mag(cf z) { return fmax(fabs((double)z.re), fabs((double)z.im)); }

// Per row, over a chunk of columns: the largest magnitude (times colscale) and the
// smallest nonzero one; the diagonal magnitude where the chunk holds it.
template <class T>
__global__ void k_eq_rows(const cx<T> *A, int64_t n, int chunk, const double *colscale, double *pmax, double *pmin,
                          double *diag) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t c0 = (int64_t)blockIdx.y * chunk, c1 = min(c0 + chunk, n);
    double mx = 0.0, mn = INFINITY;
#pragma unroll 4
    for (int64_t c = c0; c < c1; ++c) {
        const double m = eq_mag(A[c * n + r]) * (colscale ? colscale[c] : 1.0);
        mx = fmax(mx, m);
        if (m > 0.0) mn = fmin(mn, m);
        if (c == r && diag) diag[r] = m;
    }
    pmax[(int64_t)blockIdx.y * n + r] = mx;
    pmin[(int64_t)blockIdx.y * n + r] = mn;
}

__global__ void k_eq_reduce(const double *p, int parts, int64_t n, int take_min, double *out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double v = take_min ? INFINITY : 0.0;
#pragma unroll 8
    for (int q = 0; q < parts; ++q) v = take_min ? fmin(v, p[(int64_t)q * n + r]) : fmax(v, p[(int64_t)q * n + r]);
    out[r] = v;
}

// Largest magnitude (times rowscale) of each column: one workgroup per column.
template <class T>
__global__ __launch_bounds__(256) void k_eq_cols(const cx<T> *A, int64_t n, const double *rowscale, double *cmax) {
    const int64_t c = blockIdx.x;
    double mx = 0.0;
    for (int64_t r = threadIdx.x; r < n; r += 256) mx = fmax(mx, eq_mag(A[c * n + r]) * (rowscale ? rowscale[r] : 1.0));
    __shared__ double red[256];
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) cmax[c] = red[0];
}

template <class T>
__global__ void k_eq_apply(cx<T> *A, int64_t n, const double *r, const double *c) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const int64_t col = t / n, row = t - col * n;
    const double s = r[row] * c[col];  // a power of two: exact (in double, then rounded once)
    A[t] = cx<T>((T)((double)A[t].re * s), (T)((double)A[t].im * s));
}

// x[col][i] *= d[i] for `cols` packed columns of length n.
__global__ void k_scale_rows(cd *x, int64_t n, int cols, const double *d) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * cols) return;
    const double s = d[t % n];
    x[t] = cd(x[t].re * s, x[t].im * s);
}

// ---- balancing (linalg/mod.rs:247-310): power-of-two row and column factors that
// even out the magnitude sums of the max-equilibrated operator S, one sweep moving
// every column factor, then every row factor, to the power of two that minimises
// sum |s_ij| x_i y_j - sum ln x_i - sum ln y_j along it.  flags[0]: a sum was not a
// normal number (keep the max scales); flags[1]: some factor moved beyond 2.

// factor *= the power of two that puts sum * factor (given as value) / ln 2 in [1, 2).
__device__ __forceinline__ void bal_advance(double &factor, double value, int *flags) {
    const double ratio = value / 0.69314718055994530942;
    if (!(ratio >= 2.2250738585072014e-308 && ratio <= 1.7976931348623157e308)) {
        flags[0] = 1;
        return;
    }
    int e = 0;
    (void)frexp(ratio, &e);                // ratio = m 2^e, m in [0.5, 1)
    const double step = ldexp(1.0, 1 - e);  // 1 / (largest power of two <= ratio)
    factor *= step;
    if (step < 0.5 || step > 2.0) flags[1] = 1;
}

// Column pass: y_j from sum_i x_i |s_ij|; one workgroup per column.
template <class T>
__global__ __launch_bounds__(256) void k_bal_cols(const cx<T> *S, int64_t n, const double *x, double *y,
                                                  int *flags) {
    const int64_t c = blockIdx.x;
    double sum = 0.0;
    for (int64_t r = threadIdx.x; r < n; r += 256) sum += x[r] * eq_mag(S[c * n + r]);
    __shared__ double red[256];
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) bal_advance(y[c], red[0] * y[c], flags);
}

// Row pass: x_i from sum_j |s_ij| y_j.  Partial sums over column chunks (blockIdx.y),
// one thread per row (coalesced over rows), then the chunks added in order per row.
template <class T>
__global__ void k_bal_rows_part(const cx<T> *S, int64_t n, int chunk, const double *y, double *part) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t c0 = (int64_t)blockIdx.y * chunk, c1 = min(c0 + chunk, n);
    double sum = 0.0;
#pragma unroll 4
    for (int64_t c = c0; c < c1; ++c) sum += eq_mag(S[c * n + r]) * y[c];
    part[(int64_t)blockIdx.y * n + r] = sum;
}

__global__ void k_bal_rows(const double *part, int parts, int64_t n, double *x, int *flags) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double sum = 0.0;
    for (int q = 0; q < parts; ++q) sum += part[(int64_t)q * n + r];
    bal_advance(x[r], sum * x[r], flags);
}

__global__ void k_fill_f64(double *p, int64_t count, double v) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = v;
}

__global__ void k_fill_i32(int *p, int count, int v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = v;
}

constexpr int kBalanceSweeps = 64;  // dense operators take at most three (linalg/mod.rs:240)

// Balance the equilibrated S in place and fold the factors into dr, dc (device): the
// sweeps run until none moves a factor beyond 2, at most kBalanceSweeps; a non-normal
// sum or a factor beyond 2^+-256 keeps the max scales.
template <class T>
static int balance_t(cx<T> *S, int64_t n, double *dr, double *dc, hipStream_t s) {
    // column chunks of the row pass: about 2048 workgroups, at least 32 columns each
    const int64_t rblocks = (n + 255) / 256;
    const int chunk = (int)std::min<int64_t>(256, std::max<int64_t>(32, (n * rblocks + 2047) / 2048));
    const int parts = (int)((n + chunk - 1) / chunk);
    DevBuf work;
    TG_TRY(work.ensure(sizeof(double) * (2 + (size_t)parts) * (size_t)n + 256));
    double *x = work.as<double>(), *y = x + n, *part = y + n;
    int *flags = reinterpret_cast<int *>(part + (size_t)parts * n);
    const unsigned g = (unsigned)((n + 255) / 256);
    k_fill_f64<<<(unsigned)((2 * n + 255) / 256), 256, 0, s>>>(x, 2 * n, 1.0);
    int hf[2] = {0, 0};
    for (int sweep = 0; sweep < kBalanceSweeps; ++sweep) {
        k_fill_i32<<<1, 64, 0, s>>>(flags, 2, 0);
        k_bal_cols<T><<<(unsigned)n, 256, 0, s>>>(S, n, x, y, flags);
        k_bal_rows_part<T><<<dim3(g, (unsigned)parts), 256, 0, s>>>(S, n, chunk, y, part);
        k_bal_rows<<<g, 256, 0, s>>>(part, parts, n, x, flags);
        TG_HIP(hipGetLastError());
        TG_HIP(hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, s));
        TG_HIP(hipStreamSynchronize(s));
        if (hf[0] || !hf[1]) break;
    }
    if (hf[0]) return TG_OK;  // not normal: the max scales stay
    std::vector<double> h(4 * (size_t)n);
    TG_HIP(hipMemcpyAsync(h.data(), x, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s));
    TG_HIP(hipMemcpyAsync(h.data() + 2 * n, dr, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    TG_HIP(hipMemcpyAsync(h.data() + 3 * n, dc, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    const double bound = std::ldexp(1.0, 256);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 2; ++k) {
            const double fct = h[k * n + i], scale = h[(2 + k) * n + i] * fct;
            if (!(fct >= 1.0 / bound && fct <= bound) || !std::isnormal(scale)) return TG_OK;
        }
    k_eq_apply<T><<<(unsigned)((n * n + 255) / 256), 256, 0, s>>>(S, n, x, y);
    for (int64_t i = 0; i < n; ++i) {
        h[2 * n + i] *= h[i];
        h[3 * n + i] *= h[n + i];
    }
    TG_HIP(hipMemcpyAsync(dr, h.data() + 2 * n, sizeof(double) * n, hipMemcpyHostToDevice, s));
    TG_HIP(hipMemcpyAsync(dc, h.data() + 3 * n, sizeof(double) * n, hipMemcpyHostToDevice, s));
    TG_HIP(hipGetLastError());
    TG_HIP(hipStreamSynchronize(s));
    work.release();
    return TG_OK;
}

// 2^-e for v = m 2^e, m in [0.5, 1): the power of two that brings v into [0.5, 1).
static double pow2_inverse(double v) {
    int e = 0;
    (void)std::frexp(v, &e);
    return std::ldexp(1.0, -e);
}

template <class T>
static int equilibrate_t(tg_factor *f) {
    tg_cluster *c = f->cluster;
    hipStream_t s = c->stream;
    const int64_t n = f->n;
    const cx<T> *A = reinterpret_cast<const cx<T> *>(f->lu.ptr);
    // columns per row-partial: enough partials for about 2048 workgroups, at least 32
    // columns each (a 240-row operator otherwise runs one workgroup over all columns;
    // narrower chunks make the in-order reduction of the partials the slow part)
    const int64_t rblocks = (n + 255) / 256;
    const int chunk = (int)std::min<int64_t>(256, std::max<int64_t>(32, (n * rblocks + 2047) / 2048));
    const int parts = (int)((n + chunk - 1) / chunk);
    DevBuf &work = f->eqwork;
    TG_TRY(work.ensure(sizeof(double) * ((size_t)2 * parts * n + 6 * (size_t)n)));
    double *pmax = work.as<double>(), *pmin = pmax + (size_t)parts * n, *rmax = pmin + (size_t)parts * n;
    double *rmin = rmax + n, *diag = rmin + n, *cmax = diag + n, *scale = cmax + n, *second = scale + n;
    const dim3 grows((unsigned)((n + 255) / 256), (unsigned)parts);
    const unsigned g1 = (unsigned)((n + 255) / 256);
    k_eq_rows<T><<<grows, 256, 0, s>>>(A, n, chunk, nullptr, pmax, pmin, diag);
    k_eq_reduce<<<g1, 256, 0, s>>>(pmax, parts, n, 0, rmax);
    k_eq_reduce<<<g1, 256, 0, s>>>(pmin, parts, n, 1, rmin);
    k_eq_cols<T><<<(unsigned)n, 256, 0, s>>>(A, n, nullptr, cmax);
    TG_HIP(hipGetLastError());
    std::vector<double> h(4 * (size_t)n);
    TG_HIP(hipMemcpyAsync(h.data(), rmax, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    const double *hr = h.data(), *hmin = hr + n, *hd = hmin + n, *hc = hd + n;
    double maximum = 0.0, minimum = INFINITY, row_min = INFINITY, col_min = INFINITY;
    bool diag_ok = true;
    for (int64_t i = 0; i < n; ++i) {
        if (!std::isfinite(hr[i]) || !std::isfinite(hc[i])) {
            set_error("non-finite entry in the cluster operator");
            return TG_INVALID_INPUT;
        }
        maximum = std::max(maximum, hr[i]);
        minimum = std::min(minimum, hmin[i]);
        row_min = std::min(row_min, hr[i]);
        col_min = std::min(col_min, hc[i]);
    }
    for (int64_t i = 0; i < n; ++i) diag_ok = diag_ok && hd[i] / maximum >= 0.1;
    f->equilibrated = false;
    if (maximum > 0.0 && (minimum / maximum >= 0.1 || diag_ok)) return TG_OK;
    if (row_min == 0.0 || col_min == 0.0) {
        set_error("zero row or column in the cluster operator");
        return TG_SINGULAR;
    }
    if (row_min / maximum >= 0.1 && col_min / maximum >= 0.1) return TG_OK;
    // the axis with the smaller minimum first, then the other measured after it
    const bool columns_first = col_min < row_min;
    std::vector<double> first(n), other(n);
    for (int64_t i = 0; i < n; ++i) first[i] = pow2_inverse(columns_first ? hc[i] : hr[i]);
    TG_HIP(hipMemcpyAsync(scale, first.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (columns_first) {
        k_eq_rows<T><<<grows, 256, 0, s>>>(A, n, chunk, scale, pmax, pmin, nullptr);
        k_eq_reduce<<<g1, 256, 0, s>>>(pmax, parts, n, 0, second);
    } else {
        k_eq_cols<T><<<(unsigned)n, 256, 0, s>>>(A, n, scale, second);
    }
    TG_HIP(hipGetLastError());
    TG_HIP(hipMemcpyAsync(other.data(), second, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < n; ++i) {
        if (!(other[i] > 0.0)) {
            set_error("zero row or column in the cluster operator");
            return TG_SINGULAR;
        }
        other[i] = pow2_inverse(other[i]);
    }
    TG_TRY(f->eq.ensure(sizeof(double) * 2 * (size_t)n));
    double *dr = f->eq.as<double>(), *dc = dr + n;
    TG_HIP(hipMemcpyAsync(dr, (columns_first ? other : first).data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    TG_HIP(hipMemcpyAsync(dc, (columns_first ? first : other).data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    k_eq_apply<T><<<(unsigned)((n * n + 255) / 256), 256, 0, s>>>(reinterpret_cast<cx<T> *>(f->lu.ptr), n, dr, dc);
    TG_HIP(hipGetLastError());
    TG_TRY(balance_t<T>(reinterpret_cast<cx<T> *>(f->lu.ptr), n, dr, dc, s));
    TG_HIP(hipStreamSynchronize(s));
    f->equilibrated = true;
    return TG_OK;
}

__global__ void k_c128_to_c64(const cd *in, int64_t count, cf *out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) out[t] = cf((float)in[t].re, (float)in[t].im);
}

// Assemble (or, for a dense factor, copy) the operator into the factor storage,
// equilibrate and factor it, pivoting inside 64-row blocks or over whole columns.
static int dense_factor_error() {
    set_error("a factor of a caller's dense matrix has no cluster: use tg_dense_factor_solve");
    return TG_INVALID_INPUT;
}

static int factor_cluster(tg_factor *f, bool whole_column) {
    tg_cluster *c = f->cluster;
    const int64_t nn = f->n * f->n;
    if (f->dense) {
        if (f->precision == TG_STORE_C64)
            k_c128_to_c64<<<(unsigned)((nn + 255) / 256), 256, 0, c->stream>>>(f->op.as<cd>(), nn, f->lu.as<cf>());
        else
            TG_HIP(hipMemcpyAsync(f->lu.ptr, f->op.ptr, sizeof(cd) * (size_t)nn, hipMemcpyDeviceToDevice, c->stream));
        TG_HIP(hipGetLastError());
    } else {
        TG_TRY(launch_assemble_operator(c, f->precision, f->lu.ptr));
        if (f->precision == TG_STORE_C128 && f->n <= kDenseResidual) {
            const size_t bytes = sizeof(cd) * (size_t)nn;
            TG_TRY(f->op.ensure(bytes));
            TG_HIP(hipMemcpyAsync(f->op.ptr, f->lu.ptr, bytes, hipMemcpyDeviceToDevice, c->stream));
        }
    }
    TG_TRY(f->precision == TG_STORE_C64 ? equilibrate_t<float>(f) : equilibrate_t<double>(f));
    Timer::Rec rec;
    TG_TRY(c->timer.begin("lu_factor", c->stream, rec));
    const int st = whole_column ? luset_factor_gepp(f->set, c->stream) : luset_factor(f->set, c->stream);
    TG_TRY(c->timer.end("lu_factor", c->stream, rec));
    f->pivoting = whole_column ? 1 : 0;
    return st;
}

// Factor, and on a zero pivot inside the 64-row blocks factor again with whole-column
// pivoting.
static int factor_with_fallback(tg_factor *f) {
    const int st = factor_cluster(f, false);
    if (st != TG_SINGULAR) return st;
    return factor_cluster(f, true);
}

// The fallback of a stalled or singular solve: the operator factored again in c128
// with whole-column pivoting (once per factor).
static int refactor_robust(tg_factor *f) {
    tg_cluster *c = f->cluster;
    if (f->precision == TG_STORE_C64) {
        TG_HIP(hipStreamSynchronize(c->stream));
        luset_release(f->set);
        TG_TRY(f->lu.ensure(sizeof(cd) * (size_t)f->n * (size_t)f->n));
        f->precision = TG_STORE_C128;
        TG_TRY(luset_init(f->set, TG_STORE_C128, f->lu.ptr, {0}, {(int)f->n}));
    }
    f->fallback = 1;
    return factor_cluster(f, true);
}

// x <- A^{-1} x (A^{-H} x): with equilibration A = D_r^{-1} S D_c^{-1}, so
// A^{-1} = D_c S^{-1} D_r and A^{-H} = D_r S^{-H} D_c.
static int lu_solve(tg_factor *f, cd *x, int cols, int adjoint) {
    tg_cluster *c = f->cluster;
    const int64_t n = f->n;
    const double *dr = f->equilibrated ? f->eq.as<double>() : nullptr, *dc = dr ? dr + n : nullptr;
    const unsigned grid = (unsigned)((n * cols + 255) / 256);
    Timer::Rec rec;
    TG_TRY(c->timer.begin("lu_solve", c->stream, rec));
    if (dr) k_scale_rows<<<grid, 256, 0, c->stream>>>(x, n, cols, adjoint ? dc : dr);
    TG_TRY(luset_solve(f->set, c->stream, x, cols, adjoint));
    if (dr) k_scale_rows<<<grid, 256, 0, c->stream>>>(x, n, cols, adjoint ? dr : dc);
    TG_HIP(hipGetLastError());
    TG_TRY(c->timer.end("lu_solve", c->stream, rec));
    return TG_OK;
}

// C = A B for column-major complex matrices (M x K times K x N) on MFMA in the 3M
// form, one 64 x 64 tile of C per workgroup.
template <class T, bool KTAIL>
__global__ __launch_bounds__(256, 2) void k_zgemm_plain(const cx<T> *A, int64_t lda, const cx<T> *B, int64_t ldb,
                                                        cx<T> *C, int64_t ldc, int64_t M, int64_t N, int K) {
    const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
    __shared__ TileLds<T> sm;
    Tile3M<T> acc;
    zgemm_tile_3m<T, KTAIL>(A, lda, B, ldb, r0, c0, M, N, K, sm, acc);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32, rr = lane & 15;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t gr = r0 + wm + a * 16 + rr, gc = c0 + wn + b * 16 + mfma_out_row<T>(lane, v);
                const T t1 = acc.t1[a][b][v], t2 = acc.t2[a][b][v], t3 = acc.t3[a][b][v];
                if (gr < M && gc < N) C[gc * ldc + gr] = cx<T>(t1 - t2, t3 - t1 - t2);
            }
}

// Slice z of a split-K product: P_z = A[:, z ks : (z + 1) ks] B[z ks : (z + 1) ks, :]
// (the last slice ragged) into P + z pstride, one 64 x 64 tile per workgroup.  Small
// products (few tiles) run as several slices side by side instead of one long K loop
// per tile.
template <class T>
__global__ __launch_bounds__(256, 2) void k_zgemm_splitk(const cx<T> *A, int64_t lda, const cx<T> *B, int64_t ldb,
                                                         cx<T> *P, int64_t ldp, int64_t pstride, int64_t M,
                                                         int64_t N, int K, int ks) {
    const int z = blockIdx.z, k0 = z * ks, kz = min(ks, K - k0);
    if (kz <= 0) return;  // uniform over the workgroup
    const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
    __shared__ TileLds<T> sm;
    Tile3M<T> acc;
    zgemm_tile_3m<T, true>(A + (int64_t)k0 * lda, lda, B + k0, ldb, r0, c0, M, N, kz, sm, acc);
    cx<T> *C = P + (int64_t)z * pstride;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32, rr = lane & 15;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t gr = r0 + wm + a * 16 + rr, gc = c0 + wn + b * 16 + mfma_out_row<T>(lane, v);
                const T t1 = acc.t1[a][b][v], t2 = acc.t2[a][b][v], t3 = acc.t3[a][b][v];
                if (gr < M && gc < N) C[gc * ldp + gr] = cx<T>(t1 - t2, t3 - t1 - t2);
            }
}

// r = b - sum over the S slices of P (slices in order).
__global__ void k_residual_slices(const cd *b, const cd *P, int64_t count, int S, int64_t pstride, cd *r) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    cd s = b[t];
    for (int z = 0; z < S; ++z) s -= P[(int64_t)z * pstride + t];
    r[t] = s;
}

// y = A^H x for the stored FP64 operator (n x n, column-major): one workgroup per
// output row j = column j of A, reduced over its contiguous entries, `cols` columns.
__global__ __launch_bounds__(256) void k_gemv_adjoint(const cd *A, int64_t n, const cd *x, cd *y, int cols) {
    const int64_t j = blockIdx.x;
    const cd *a = A + j * n;
    __shared__ double red[2][256];
    for (int cc = 0; cc < cols; ++cc) {
        cd acc(0.0);
        const cd *xc = x + (int64_t)cc * n;
        for (int64_t i = threadIdx.x; i < n; i += 256) acc += conj(a[i]) * xc[i];
        red[0][threadIdx.x] = acc.re;
        red[1][threadIdx.x] = acc.im;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) {
                red[0][threadIdx.x] += red[0][threadIdx.x + w];
                red[1][threadIdx.x] += red[1][threadIdx.x + w];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) y[(int64_t)cc * n + j] = cd(red[0][0], red[1][0]);
        __syncthreads();
    }
}

static int factor_solve_once(tg_factor *f, const cd *b, cd *x, int cols, int adjoint, double rtol,
                             std::vector<tg_convergence> &conv);

// x = A^{-1} b (or A^{-H} b) in the split order with FP64 iterative refinement on the
// matrix-free FP64 operator, until |b - A x| <= rtol |b| for every column.  A solve
// that stalls or meets a non-finite residual refactors the operator once in c128 with
// whole-column pivoting (reported by tg_factor_info) and is repeated.
int factor_solve_split(tg_factor *f, const cd *b, cd *x, int cols, int adjoint, double rtol,
                       std::vector<tg_convergence> &conv) {
    int st = factor_solve_once(f, b, x, cols, adjoint, rtol, conv);
    if ((st != TG_NOT_CONVERGED && st != TG_SINGULAR) || f->fallback) return st;
    TG_TRY(refactor_robust(f));
    return factor_solve_once(f, b, x, cols, adjoint, rtol, conv);
}

static int factor_solve_once(tg_factor *f, const cd *b, cd *x, int cols, int adjoint, double rtol,
                             std::vector<tg_convergence> &conv) {
    tg_cluster *c = f->cluster;
    const int64_t count = (int64_t)cols * c->n;
    // the forward residual with the stored operator: split-K slices so that about 128
    // tiles run at once (at most 8 slices of a multiple of 16 columns)
    const int64_t n = f->n;
    int S = 1, ks = (int)n;
    if (f->op.ptr && !adjoint) {
        const int64_t tiles = ((cols + 63) / 64) * ((n + 63) / 64);
        S = (int)std::min<int64_t>(8, std::max<int64_t>(1, 128 / tiles));
        ks = (int)(((n + S - 1) / S + 15) / 16 * 16);
        S = (int)((n + ks - 1) / ks);
    }
    DevBuf tmp;
    TG_TRY(tmp.ensure(sizeof(cd) * count * (1 + std::max(S, 1))));
    cd *r = tmp.as<cd>(), *ax = r + count;
    std::vector<double> bn(cols), rn(cols);
    int st = TG_OK;
    conv.assign(cols, tg_convergence{0, 0, 0.0, 0.0});
    // with equilibration the certificate is the residual of the scaled system:
    // D_r (b - A x) against D_r b (D_c for the adjoint)
    const double *dnorm = f->equilibrated ? f->eq.as<double>() + (adjoint ? f->n : 0) : nullptr;
    const unsigned sgrid = (unsigned)((count + 255) / 256);
    TG_TRY(f->norms.ensure(sizeof(cd) * 2 * (size_t)cols));
    cd *bsq = f->norms.as<cd>(), *rsq = bsq + cols;
    // squared norms into pinned host memory, read after the next synchronisation
    auto norms = [&](const cd *v, cd *host) -> int {
        if (dnorm) {
            TG_HIP(hipMemcpyAsync(ax, v, sizeof(cd) * count, hipMemcpyDeviceToDevice, c->stream));
            k_scale_rows<<<sgrid, 256, 0, c->stream>>>(ax, c->n, cols, dnorm);
            TG_HIP(hipGetLastError());
            v = ax;
        }
        return vec_dots_dev(c, v, 0, 1, v, c->n, cols, c->n, c->n, nullptr, nullptr, host);
    };
    // a NaN stays NaN (a clamp would turn it into 0 and pass a non-finite vector as zero)
    auto norm = [](cd d) { return d.re > 0.0 ? sqrt(d.re) : d.re == d.re ? 0.0 : d.re; };
    do {
        if ((st = norms(b, bsq))) break;
        if (hipMemcpyAsync(x, b, sizeof(cd) * count, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
            set_error("copy of the right-hand side failed");
            st = TG_DEVICE_ERROR;
            break;
        }
        if ((st = lu_solve(f, x, cols, adjoint))) break;
        int refinements = 0;
        const int max_refine = f->precision == TG_STORE_C64 ? 12 : 4;
        while (true) {
            if (f->op.ptr && !adjoint) {
                // r = b - A x with the stored FP64 operator (n x n times n x cols), the
                // product in S split-K slices (ax .. ax + S count)
                const dim3 grid((unsigned)((cols + 63) / 64), (unsigned)((n + 63) / 64), (unsigned)S);
                k_zgemm_splitk<double><<<grid, 256, 0, c->stream>>>(f->op.as<cd>(), n, x, n, ax, n, count, n, cols,
                                                                    (int)n, ks);
                k_residual_slices<<<(unsigned)((count + 255) / 256), 256, 0, c->stream>>>(b, ax, count, S, count, r);
                if (hipGetLastError() != hipSuccess) {
                    set_error("launch of the residual GEMM failed");
                    st = TG_DEVICE_ERROR;
                    break;
                }
            } else if (f->op.ptr && adjoint) {
                k_gemv_adjoint<<<(unsigned)f->n, 256, 0, c->stream>>>(f->op.as<cd>(), f->n, x, ax, cols);
                if (hipGetLastError() != hipSuccess) {
                    set_error("launch of the adjoint residual failed");
                    st = TG_DEVICE_ERROR;
                    break;
                }
            } else if ((st = apply_operator(c, x, ax, cols, adjoint, TG_OP_MATRIX_FREE_F64))) {
                break;
            }
            if (!(f->op.ptr && !adjoint)) {
                k_residual_slices<<<(unsigned)((count + 255) / 256), 256, 0, c->stream>>>(b, ax, count, 1, count, r);
                if (hipGetLastError() != hipSuccess) {
                    set_error("launch of the residual failed");
                    st = TG_DEVICE_ERROR;
                    break;
                }
            }
            if ((st = norms(r, rsq))) break;
            if (hipStreamSynchronize(c->stream) != hipSuccess) {
                set_error("residual norms of the LU solve failed");
                st = TG_DEVICE_ERROR;
                break;
            }
            for (int k = 0; k < cols; ++k) {
                if (refinements == 0) bn[k] = norm(bsq[k]);
                rn[k] = norm(rsq[k]);
            }
            bool done = true;
            for (int k = 0; k < cols; ++k) {
                conv[k].residual_norm = rn[k];
                conv[k].rhs_norm = bn[k];
                conv[k].refinements = refinements;
                if (!std::isfinite(rn[k])) {
                    set_error("non-finite residual after the LU solve (numerically singular operator)");
                    st = TG_SINGULAR;
                }
                if (rn[k] > rtol * bn[k]) done = false;
            }
            if (st || done) break;
            if (refinements >= max_refine) {
                set_error("LU refinement did not reach rtol %g after %d steps (relative residual %e)", rtol,
                          refinements, rn[0] / (bn[0] > 0 ? bn[0] : 1.0));
                st = TG_NOT_CONVERGED;
                break;
            }
            if ((st = lu_solve(f, r, cols, adjoint))) break;
            if ((st = vec_axpby(c->stream, cd(1.0), r, cd(1.0), x, count))) break;
            ++refinements;
        }
    } while (0);
    if (st == TG_OK) {
        // the last residual norms were read on the host: the stream is idle
        CleanReleases clean;
        tmp.release();
    } else {
        tmp.release();
    }
    return st;
}

// Gradient body shared with the GMRES pullback (tg_adjoint.hip), here with the
// adjoint solve by the factors.
int pullback_with(tg_cluster *c, const tg_c128 *incident, const tg_c128 *x, const tg_c128 *cotangent, int cols,
                  const AdjointSolve &adjoint_solve, double *grad_k0, double *grad_radii, tg_c128 *grad_eps,
                  double *grad_positions, tg_c128 *incident_grad, tg_convergence *conv_host);

}  // namespace tg

using namespace tg;

extern "C" {

int tg_factor_create(tg_cluster *c, int precision, tg_factor **out) {
    if (!c || !out || (precision != TG_STORE_C64 && precision != TG_STORE_C128)) {
        set_error("null argument or precision other than TG_STORE_C64 / TG_STORE_C128");
        return TG_INVALID_INPUT;
    }
    *out = nullptr;
    tg_factor *f = new tg_factor;
    f->cluster = c;
    f->precision = precision;
    f->n = c->n;
    const size_t esz = precision == TG_STORE_C64 ? sizeof(cf) : sizeof(cd);
    int st = f->lu.ensure(esz * (size_t)c->n * (size_t)c->n);
    if (!st) st = luset_init(f->set, precision, f->lu.ptr, {0}, {(int)c->n});
    if (!st) st = factor_with_fallback(f);
    if (st) {
        (void)hipStreamSynchronize(c->stream);
        f->lu.release();
        f->eq.release();
        f->op.release();
        luset_release(f->set);
        delete f;
        return st;
    }
    *out = f;
    return TG_OK;
}

int tg_factor_destroy(tg_factor *f) {
    if (!f) return TG_OK;
    (void)hipStreamSynchronize(f->cluster->stream);
    // the panel stream's work joined the cluster's stream: the blocks are idle
    CleanReleases clean;
    f->eqwork.release();
    f->lu.release();
    f->eq.release();
    f->op.release();
    f->norms.release();
    luset_release(f->set);
    if (f->dense) {
        for (auto &w : f->cluster->ws) w.release();
        delete f->cluster;
    }
    delete f;
    return TG_OK;
}

int tg_dense_factor_create(int64_t n, const tg_c128 *a, int precision, void *stream, tg_factor **out) {
    if (!out || !a || n < 1 || n > INT32_MAX || (precision != TG_STORE_C64 && precision != TG_STORE_C128)) {
        set_error("null argument, n outside 1..2^31-1 or precision other than TG_STORE_C64 / TG_STORE_C128");
        return TG_INVALID_INPUT;
    }
    *out = nullptr;
    tg_factor *f = new tg_factor;
    f->cluster = new tg_cluster();
    f->cluster->n = n;
    f->cluster->stream = (hipStream_t)stream;
    f->dense = true;
    f->precision = precision;
    f->n = n;
    const size_t nn = (size_t)n * (size_t)n;
    const size_t esz = precision == TG_STORE_C64 ? sizeof(cf) : sizeof(cd);
    int st = f->op.ensure(sizeof(cd) * nn);
    if (!st && hipMemcpyAsync(f->op.ptr, a, sizeof(cd) * nn, hipMemcpyDeviceToDevice, f->cluster->stream) != hipSuccess) {
        set_error("copy of the matrix failed");
        st = TG_DEVICE_ERROR;
    }
    if (!st) st = f->lu.ensure(esz * nn);
    if (!st) st = luset_init(f->set, precision, f->lu.ptr, {0}, {(int)n});
    if (!st) st = factor_with_fallback(f);
    if (st) {
        tg_factor_destroy(f);
        return st;
    }
    *out = f;
    return TG_OK;
}

int tg_dense_factor_solve(tg_factor *f, const tg_c128 *b, int cols, double rtol, int adjoint, tg_c128 *x,
                          tg_convergence *conv_host) {
    if (!f || !b || !x || cols < 1 || !(rtol > 0.0)) {
        set_error("null argument, no columns or non-positive rtol");
        return TG_INVALID_INPUT;
    }
    std::vector<tg_convergence> conv;
    int st = factor_solve_split(f, (const cd *)b, (cd *)x, cols, adjoint, rtol, conv);
    if (hipStreamSynchronize(f->cluster->stream) != hipSuccess && !st) {
        set_error("device failure in the LU solve");
        st = TG_DEVICE_ERROR;
    }
    if (conv_host)
        for (int k = 0; k < cols && k < (int)conv.size(); ++k) conv_host[k] = conv[k];
    return st;
}

int tg_factor_info(const tg_factor *f, int *precision, int *pivoting, int *fallback) {
    if (!f) {
        set_error("null factor");
        return TG_INVALID_INPUT;
    }
    if (precision) *precision = f->precision;
    if (pivoting) *pivoting = f->pivoting;
    if (fallback) *fallback = f->fallback;
    return TG_OK;
}

int tg_factor_solve(tg_factor *f, const tg_c128 *incident, int cols, double rtol, int adjoint, tg_c128 *x,
                    tg_convergence *conv_host) {
    if (f && f->dense) return dense_factor_error();
    if (!f || !incident || !x || cols < 1 || !(rtol > 0.0)) {
        set_error("null argument, no columns or non-positive rtol");
        return TG_INVALID_INPUT;
    }
    tg_cluster *c = f->cluster;
    const int64_t count = (int64_t)cols * c->n;
    DevBuf io;
    TG_TRY(io.ensure(sizeof(cd) * count * 3));
    cd *a = io.as<cd>(), *b = a + count, *xs = b + count;
    std::vector<tg_convergence> conv;
    int st = ref_to_split(c, (const cd *)incident, a, cols);
    // rhs = T a for the forward solve (InteractionFactor::solve); the adjoint solve
    // takes its right-hand side as it is (Lu::solve_adjoint_in_place)
    if (!st) {
        if (adjoint) {
            if (hipMemcpyAsync(b, a, sizeof(cd) * count, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
                set_error("copy of the right-hand side failed");
                st = TG_DEVICE_ERROR;
            }
        } else {
            st = local_apply(c, a, b, cols, 0, nullptr);
        }
    }
    if (!st) st = factor_solve_split(f, b, xs, cols, adjoint, rtol, conv);
    if (!st) st = split_to_ref(c, xs, (cd *)x, cols);
    const bool idle = hipStreamSynchronize(c->stream) == hipSuccess;
    if (!st && !idle) {
        set_error("device failure in the LU solve");
        st = TG_DEVICE_ERROR;
    }
    if (conv_host)
        for (int k = 0; k < cols && k < (int)conv.size(); ++k) conv_host[k] = conv[k];
    if (idle) {
        CleanReleases clean;
        io.release();
    } else {
        io.release();
    }
    return st;
}

int tg_factor_equilibrated(const tg_factor *f) { return f && f->equilibrated ? 1 : 0; }

int tg_factor_global_tmatrix(tg_factor *f, int lmax_global, double rtol, tg_c128 *out, tg_convergence *conv_host) {
    if (f && f->dense) return dense_factor_error();
    if (!f || !out || !(rtol > 0.0)) {
        set_error("null argument or non-positive rtol");
        return TG_INVALID_INPUT;
    }
    tg_cluster *c = f->cluster;
    HostPlan pin, pout;  // R(p_j <- 0): lmax <- lmax_global; R(0 <- p_i): lmax_global <- lmax
    if (lmax_global < 1 || build_plan_rect(c->lmax, lmax_global, pin) || build_plan_rect(lmax_global, c->lmax, pout)) {
        set_error("global lmax must be >= 1 with lmax + global lmax <= %d", 2 * kMaxLmax);
        return TG_INVALID_INPUT;
    }
    const int N = c->N, M = c->M, Mg = 2 * lmax_global * (lmax_global + 2);
    const int64_t n = c->n, MMg = (int64_t)M * Mg;
    std::vector<double> disp(6 * (size_t)N);
    for (int j = 0; j < N; ++j)
        for (int a = 0; a < 3; ++a) {
            disp[3 * j + a] = c->positions[3 * j + a];
            disp[3 * (N + j) + a] = -c->positions[3 * j + a];
        }
    DevBuf buf, pbin, pbout;
    RectPlan rin_plan{}, rout_plan{};
    TG_TRY(upload_rect_plan(pin, pbin, rin_plan, c->stream));
    TG_TRY(upload_rect_plan(pout, pbout, rout_plan, c->stream));
    TG_TRY(buf.ensure(sizeof(cd) * (2 * N * MMg + 2 * n * Mg) + sizeof(double) * disp.size()));
    cd *blocks = buf.as<cd>(), *rin = blocks + 2 * N * MMg, *z = rin + n * Mg;
    double *ddisp = reinterpret_cast<double *>(z + n * Mg);
    int st = TG_OK;
    do {
        if (hipMemcpyAsync(ddisp, disp.data(), sizeof(double) * disp.size(), hipMemcpyHostToDevice, c->stream) !=
            hipSuccess) {
            set_error("copy of the displacements failed");
            st = TG_DEVICE_ERROR;
            break;
        }
        // incident side: N blocks M x Mg; outgoing side: N blocks Mg x M
        if ((st = launch_translation_rect(rin_plan, c->k0, N, ddisp, 1, blocks, c->stream))) break;
        if ((st = launch_translation_rect(rout_plan, c->k0, N, ddisp + 3 * N, 1, blocks + N * MMg, c->stream))) break;
        // Rin (n x Mg): rows (j, a), columns g; block j holds R(p_j <- 0)[a][g] at g M + a
        for (int j = 0; j < N && !st; ++j)
            if (hipMemcpy2DAsync(rin + (int64_t)j * M, sizeof(cd) * n, blocks + j * MMg, sizeof(cd) * M, sizeof(cd) * M,
                                 Mg, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
                set_error("copy of the translation blocks failed");
                st = TG_DEVICE_ERROR;
            }
        if (st) break;
        // Z = (I - T C)^{-1} T Rin, then T_g = Rout Z, Rout = [R(0 <- p_1) ... R(0 <- p_N)] (Mg x n)
        if ((st = tg_factor_solve(f, reinterpret_cast<const tg_c128 *>(rin), Mg, rtol, 0,
                                  reinterpret_cast<tg_c128 *>(z), conv_host)))
            break;
        const cd *rout = blocks + N * MMg;
        const dim3 grid((unsigned)((Mg + 63) / 64), (unsigned)((Mg + 63) / 64));
        if (n % 16)
            k_zgemm_plain<double, true><<<grid, 256, 0, c->stream>>>(rout, Mg, z, n, (cd *)out, Mg, Mg, Mg, (int)n);
        else
            k_zgemm_plain<double, false><<<grid, 256, 0, c->stream>>>(rout, Mg, z, n, (cd *)out, Mg, Mg, Mg, (int)n);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
            set_error("device failure in the global T-matrix");
            st = TG_DEVICE_ERROR;
        }
    } while (0);
    (void)hipStreamSynchronize(c->stream);
    buf.release();
    pbin.release();
    pbout.release();
    return st;
}

int tg_factor_pullback(tg_factor *f, const tg_c128 *incident, const tg_c128 *x, const tg_c128 *cotangent, int cols,
                       double rtol, double *grad_k0, double *grad_radii, tg_c128 *grad_eps, double *grad_positions,
                       tg_c128 *incident_grad, tg_convergence *conv_host) {
    if (f && f->dense) return dense_factor_error();
    if (!f || !(rtol > 0.0)) {
        set_error("null factor or non-positive rtol");
        return TG_INVALID_INPUT;
    }
    AdjointSolve solve = [f, rtol](const cd *g, cd *y, int k, std::vector<tg_convergence> &conv) {
        return factor_solve_split(f, g, y, k, 1, rtol, conv);
    };
    return pullback_with(f->cluster, incident, x, cotangent, cols, solve, grad_k0, grad_radii, grad_eps,
                         grad_positions, incident_grad, conv_host);
}

int tg_factor_pushforward(tg_factor *f, const tg_c128 *incident, const tg_c128 *x, int cols, double rtol,
                          double tangent_k0, const double *tangent_radii, const tg_c128 *tangent_eps,
                          const double *tangent_positions, const tg_c128 *tangent_incident, tg_c128 *dx,
                          tg_convergence *conv_host) {
    if (f && f->dense) return dense_factor_error();
    if (!f || !(rtol > 0.0)) {
        set_error("null factor or non-positive rtol");
        return TG_INVALID_INPUT;
    }
    TangentSolve solve = [f, rtol](const cd *b, cd *y, int k, std::vector<tg_convergence> &conv) {
        return factor_solve_split(f, b, y, k, 0, rtol, conv);
    };
    return pushforward_with(f->cluster, incident, x, cols, tangent_k0, tangent_radii, tangent_eps,
                            tangent_positions, nullptr, true, tangent_incident, solve, dx, conv_host);
}

int tg_factor_pushforward_local(tg_factor *f, const tg_c128 *incident, const tg_c128 *x, int cols, double rtol,
                                double tangent_k0, const double *tangent_positions, const tg_c128 *tangent_local,
                                const tg_c128 *tangent_incident, tg_c128 *dx, tg_convergence *conv_host) {
    if (f && f->dense) return dense_factor_error();
    if (!f || !(rtol > 0.0)) {
        set_error("null factor or non-positive rtol");
        return TG_INVALID_INPUT;
    }
    TangentSolve solve = [f, rtol](const cd *b, cd *y, int k, std::vector<tg_convergence> &conv) {
        return factor_solve_split(f, b, y, k, 0, rtol, conv);
    };
    return pushforward_with(f->cluster, incident, x, cols, tangent_k0, nullptr, nullptr, tangent_positions,
                            tangent_local, false, tangent_incident, solve, dx, conv_host);
}

int tg_factor_pullback_local(tg_factor *f, const tg_c128 *incident, const tg_c128 *x, const tg_c128 *cotangent,
                             int cols, double rtol, double *grad_k0, double *grad_positions, tg_c128 *grad_local,
                             tg_c128 *incident_grad, tg_convergence *conv_host) {
    if (f && f->dense) return dense_factor_error();
    if (!f || !(rtol > 0.0)) {
        set_error("null factor or non-positive rtol");
        return TG_INVALID_INPUT;
    }
    TangentSolve solve = [f, rtol](const cd *g, cd *y, int k, std::vector<tg_convergence> &conv) {
        return factor_solve_split(f, g, y, k, 1, rtol, conv);
    };
    return pullback_local_with(f->cluster, incident, x, cotangent, cols, solve, grad_k0, grad_positions, grad_local,
                               incident_grad, conv_host);
}

}  // extern "C"
