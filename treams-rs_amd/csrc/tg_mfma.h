// Complex GEMM tiles in the 3M form on gfx950 MFMA (v_mfma_f32_16x16x4f32 for c64,
// v_mfma_f64_16x16x4f64 for c128), shared by the LU trailing update and the
// matrix-free coupling.
#pragma once
#include "tg_complex.h"

namespace tg {

typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef double f64x4v __attribute__((ext_vector_type(4)));

template <class T> struct mfma_acc;
template <> struct mfma_acc<float> { using V = f32x4v; };
template <> struct mfma_acc<double> { using V = f64x4v; };

// D[16x16] += A[16x4] B[4x16]: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16].
// The result D[i][j] sits in lane j + 16 (i / 4), element i % 4 for f32, and in lane
// j + 16 (i % 4), element i / 4 for f64.
__device__ __forceinline__ f32x4v mfma16x16x4(float a, float b, f32x4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4v mfma16x16x4(double a, double b, f64x4v c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
template <class T>
__device__ __forceinline__ int mfma_out_row(int lane, int v) {
    return sizeof(T) == 8 ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v;
}

// Plain complex value for LDS arrays (no constructors).
template <class T>
struct lcx {
    T re, im;
    __device__ __forceinline__ cx<T> get() const { return cx<T>(re, im); }
    __device__ __forceinline__ void set(cx<T> v) { re = v.re; im = v.im; }
};

// LDS staging of one 16-wide K slice: A tile BM x 16, B tile 16 x BN, split re / im.
template <class T, int BM = 64, int BN = 64>
struct TileLds {
    T Ar[BM][17], Ai[BM][17], Br[16][BN + 1], Bi[16][BN + 1];
};

// Accumulators of a BM x BN output tile computed by 4 waves of 32 x 32 (2 x 2 16x16
// MFMA tiles each; BN / 32 waves across, BM / 32 down), for the transposed product
// (B^T A^T) so that lanes run along the output rows: element v of lane l of tile
// (a, b) is row wm + 16 a + l % 16, column wn + 16 b + mfma_out_row(l, v), with
// wm = 32 (wave / (BN / 32)), wn = 32 (wave % (BN / 32)).
// Re(A B) = t1 - t2, Im(A B) = t3 - t1 - t2.
template <class T>
struct Tile3M {
    typename mfma_acc<T>::V t1[2][2], t2[2][2], t3[2][2];
};

// t += A[r0 : r0 + BM, 0 : K] B[0 : K, cols] for column-major A (lda, elements of type
// TA converted to T on load); column c of B
// (c = c0 .. c0 + 63) starts at bcol(c), which the caller keeps a valid address for
// every c (clamped).  Rows >= M are read clamped: they only feed output rows that
// the caller does not store.  With KTAIL, k >= K is read clamped and zeroed (K not a
// multiple of 16).  No load is conditional; the next K slice is loaded while the
// current one is multiplied.  256 threads.
template <class T, bool KTAIL, int BM, int BN, class TA, class BCol>
__device__ __forceinline__ void zgemm_tile_3m_cols(const cx<TA> *A, int64_t lda, BCol bcol, int64_t r0, int64_t c0,
                                                   int64_t M, int K, TileLds<T, BM, BN> &sm, Tile3M<T> &acc) {
    static_assert(BM * BN == 64 * 64 && BM % 32 == 0 && BN % 32 == 0, "4 waves of 32 x 32");
    constexpr int WC = BN / 32, NA = BM / 16, NB = BN / 16, KA = 256 / BM;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave / WC) * 32, wn = (wave % WC) * 32;
    const int rr = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc.t1[a][b][v] = acc.t2[a][b][v] = acc.t3[a][b][v] = (T)0;
    // slice loads: A (BM rows x 16) at (m = tid % BM, k = tid / BM + KA j), j < NA,
    // B (16 x BN columns) at (k = tid & 15, column = (tid >> 4) + 16 j), j < NB
    const int lm = tid % BM, lk = tid / BM, uk = tid & 15, uc = tid >> 4;
    const int64_t arow = min(r0 + lm, M - 1);
    const cx<T> *bptr[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) bptr[j] = bcol(c0 + uc + 16 * j);
    cx<T> areg[NA], breg[NB];
    auto load_slice = [&](int kk) {
#pragma unroll
        for (int j = 0; j < NA; ++j) {
            const int ka = kk + lk + KA * j;
            // A may be stored narrower than the compute type (c64 factors, FP64 vectors)
            if (KTAIL) {
                const cx<TA> av = A[(int64_t)min(ka, K - 1) * lda + arow];
                areg[j] = ka < K ? cx<T>((T)av.re, (T)av.im) : cx<T>();
            } else {
                const cx<TA> av = A[(int64_t)ka * lda + arow];
                areg[j] = cx<T>((T)av.re, (T)av.im);
            }
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int kb = kk + uk;
            if (KTAIL) {
                const cx<T> bv = bptr[j][min(kb, K - 1)];
                breg[j] = kb < K ? bv : cx<T>();
            } else {
                breg[j] = bptr[j][kb];
            }
        }
    };
    load_slice(0);
    for (int kk = 0; kk < K; kk += 16) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NA; ++j) {
            sm.Ar[lm][lk + KA * j] = areg[j].re;
            sm.Ai[lm][lk + KA * j] = areg[j].im;
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            sm.Br[uk][uc + 16 * j] = breg[j].re;
            sm.Bi[uk][uc + 16 * j] = breg[j].im;
        }
        __syncthreads();
        if (kk + 16 < K) load_slice(kk + 16);
#pragma unroll
        for (int k4 = 0; k4 < 16; k4 += 4) {
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const T br = sm.Br[k4 + kq][wn + b * 16 + rr], bi = sm.Bi[k4 + kq][wn + b * 16 + rr];
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const T ar = sm.Ar[wm + a * 16 + rr][k4 + kq], ai = sm.Ai[wm + a * 16 + rr][k4 + kq];
                    acc.t1[a][b] = mfma16x16x4(br, ar, acc.t1[a][b]);
                    acc.t2[a][b] = mfma16x16x4(bi, ai, acc.t2[a][b]);
                    acc.t3[a][b] = mfma16x16x4(br + bi, ar + ai, acc.t3[a][b]);
                }
            }
        }
    }
}

// t += A[r0 : r0 + 64, 0 : K] B[0 : K, c0 : c0 + 64] for column-major A (lda) and B
// (ldb); columns >= N are read clamped.
template <class T, bool KTAIL>
__device__ __forceinline__ void zgemm_tile_3m(const cx<T> *A, int64_t lda, const cx<T> *B, int64_t ldb, int64_t r0,
                                              int64_t c0, int64_t M, int64_t N, int K, TileLds<T> &sm,
                                              Tile3M<T> &acc) {
    zgemm_tile_3m_cols<T, KTAIL, 64, 64>(
        A, lda, [=](int64_t c) { return B + min(c, N - 1) * ldb; }, r0, c0, M, K, sm, acc);
}

}  // namespace tg
