// Matrix-free coupling C x on FP16 MFMA with two-term splits: the c64-precision
// operator of the inner solves.
//
// The pair-table GEMMs U_t = TAB_t X (harmonic index t, subset columns as in the FP32
// and FP64 forms) run on v_mfma_f32_16x16x32_f16.  Every real operand value v is held
// as hi + lo, hi = fp16(v 2^s), lo = fp16(v 2^s - hi), with a power-of-two shift s per
// table row (t, i) and per source column q that brings the row's or column's largest
// magnitude to [2^14, 2^15).  Each real product takes three MFMAs (hi hi, hi lo,
// lo hi), so a value keeps 22 significant bits while it lies within 2^17 of its row's
// or column's largest (beyond that the error is below 2^-38 of the largest); the
// products are exact in FP32 and accumulate in FP32, as in the c64 form.  The shifts
// are undone exactly (ldexp) when a tile is stored.
//
// The complex product is the 4M form (Re = Tr Xr - Ti Xi, Im = Tr Xi + Ti Xr): four
// operand fragments per side and twelve MFMAs per 16 x 16 x 32 step, against six
// fragments and nine MFMAs for 3M, which would also need split sums Tr + Ti.
//
// Layouts (K = sources j padded to Np = 32 ceil(N / 32) with zeros):
//   tables  TABH[t][i / 16][j / 32][c][lane][8]: one MFMA operand fragment per (c,
//           lane), lane = (i % 16) + 16 ((j % 32) / 8), element j % 8; components c =
//           re hi, re lo, im hi, im lo.  T Np^2 8 bytes (the size of the c64 tables).
//   shifts  row exponents E[t][i] (|TAB_t[i][j]| < 2^E), shift 15 - E.
//   sources XH[q][c][Np] halves with the column shift cshift[q].
#include "tg_internal.h"
#include "tg_mfma.h"

namespace tg {

typedef _Float16 h8v __attribute__((ext_vector_type(8)));

constexpr int kNoExp = -100000;  // row or column of zeros

__device__ __forceinline__ int mag_exp(double m) {
    int e = 0;
    (void)frexp(m, &e);
    return m > 0.0 ? e : kNoExp;
}
__device__ __forceinline__ int shift_of(int e) { return e == kNoExp ? 0 : 15 - e; }

// hi and lo halves of v 2^s
struct half2s { _Float16 hi, lo; };
__device__ __forceinline__ half2s split16(double v, int s) {
    const double w = ldexp(v, s);
    const _Float16 hi = (_Float16)(float)w;
    return {hi, (_Float16)(float)(w - (double)(float)hi)};
}

__global__ void k_fill_int(int *p, int64_t count, int v) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) p[t] = v;
}

// E[t][i] = max over j in chunk blockIdx.z of the exponent of max(|re|, |im|) of
// TAB[t][j][i]; grid (ceil(N / 256), T, chunks).
__global__ __launch_bounds__(256) void k_split_rowexp(const cd *TAB, int N, int Np, int jchunk, int *E) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int t = blockIdx.y;
    if (i >= N) return;
    const int j0 = blockIdx.z * jchunk, j1 = min(N, j0 + jchunk);
    const cd *p = TAB + ((int64_t)t * N) * N + i;
    double m = 0.0;
    for (int j = j0; j < j1; ++j) {
        const cd v = p[(int64_t)j * N];
        m = fmax(m, fmax(fabs(v.re), fabs(v.im)));
    }
    const int e = mag_exp(m);
    if (e != kNoExp) atomicMax(E + (int64_t)t * Np + i, e);
}

// The same row maxima from the int16 exponents of the table entries (written with the
// FP64 tables): an eighth of the bytes.
__global__ __launch_bounds__(256) void k_split_rowexp16(const int16_t *ex, int N, int Np, int jchunk, int *E) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int t = blockIdx.y;
    if (i >= N) return;
    const int j0 = blockIdx.z * jchunk, j1 = min(N, j0 + jchunk);
    const int16_t *p = ex + ((int64_t)t * N) * N + i;
    int m = -32768;
#pragma unroll 8
    for (int j = j0; j < j1; ++j) m = max(m, (int)p[(int64_t)j * N]);
    if (m != -32768) atomicMax(E + (int64_t)t * Np + i, m);
}

// TABH from the FP64 tables: one thread per (t, i block, j block, lane).
__global__ __launch_bounds__(256) void k_split_tables(const cd *TAB, const int *E, int N, int Np, int T,
                                                      h8v *TABH) {
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int NB16 = Np / 16, NB32 = Np / 32;
    const int64_t nfrag = (int64_t)T * NB16 * NB32;
    const int64_t frag = gid >> 6;
    if (frag >= nfrag) return;
    const int lane = (int)(gid & 63);
    const int jb = (int)(frag % NB32);
    const int64_t rest = frag / NB32;
    const int ib = (int)(rest % NB16), t = (int)(rest / NB16);
    const int i = ib * 16 + (lane & 15), j0 = jb * 32 + 8 * (lane >> 4);
    const int s = i < N ? shift_of(E[(int64_t)t * Np + i]) : 0;
    h8v rh, rl, ih, il;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int j = j0 + e;
        cd v(0.0);
        if (i < N && j < N) v = TAB[((int64_t)t * N + j) * N + i];
        const half2s pr = split16(v.re, s), pi = split16(v.im, s);
        rh[e] = pr.hi;
        rl[e] = pr.lo;
        ih[e] = pi.hi;
        il[e] = pi.lo;
    }
    h8v *out = TABH + frag * 256 + lane;
    out[0] = rh;
    out[64] = rl;
    out[128] = ih;
    out[192] = il;
}

// XH[q][c][j] and cshift[q] for q = (column cc, helicity h, mode a) of x (split
// layout); one workgroup per q.
__global__ __launch_bounds__(256) void k_split_sources(const cd *x, int64_t n, int64_t nh, int N, int Np, int H,
                                                       _Float16 *XH, int *cshift) {
    const int q = blockIdx.x;
    const int cc = q / (2 * H), rem = q - cc * 2 * H, h = rem / H, a = rem - h * H;
    const cd *xq = x + (int64_t)cc * n + (int64_t)h * nh + a;
    double m = 0.0;
    for (int j = threadIdx.x; j < N; j += 256) {
        const cd v = xq[(int64_t)j * H];
        m = fmax(m, fmax(fabs(v.re), fabs(v.im)));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    __shared__ double wm[4];
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    m = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
    const int s = shift_of(mag_exp(m));
    if (threadIdx.x == 0) cshift[q] = s;
    _Float16 *o = XH + (int64_t)q * 4 * Np;
    for (int j = threadIdx.x; j < Np; j += 256) {
        const cd v = j < N ? xq[(int64_t)j * H] : cd(0.0);
        const half2s pr = split16(v.re, s), pi = split16(v.im, s);
        o[j] = pr.hi;
        o[Np + j] = pr.lo;
        o[2 * Np + j] = pi.hi;
        o[3 * Np + j] = pi.lo;
    }
}

__device__ __forceinline__ f32x4v mfma_h(h8v a, h8v b, f32x4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// U[g][gpos[soff[t] + r]][i] = sum_j TAB_t[i][j] X[g H + smode[soff[t] + r]][j] (c64
// out; columns in the contraction's group order).
// One 256 x 64 tile (rows i, subset columns) per workgroup of 8 waves; wave w owns
// rows 32 w .. 32 w + 31 of the tile and all 64 columns (2 x 4 MFMA tiles), computed
// transposed (columns as the MFMA's M, rows i as its N) so that lanes run along i in
// the stores.  The table fragments are loaded straight from global memory (the layout
// above is fragment-native), the next K step's while the current one multiplies; the
// source tile, shared by the eight waves, goes through LDS (double-buffered, one
// barrier per K step).  Tile order and the XCD remap as in the FP32 form, with
// 64-wide column tiles (tiles[t]) and 256-row tiles.
// LO = false multiplies the hi halves only (11 significant bits per value, four MFMAs
// per step instead of twelve, half the table fragments): the relaxed applies of an
// inner solve whose residual is already small (coupling_tables_split16).
template <bool LO>
__global__ __launch_bounds__(512, 1) void k_zgemm_subsets_split16(const h8v *TABH, const _Float16 *XH,
                                                                  const int *E, const int *cshift, cf *U, int N,
                                                                  int Np, int H, int S, int G, int T_,
                                                                  const int32_t *tiles, const int32_t *soff,
                                                                  const int32_t *smode, const int32_t *gpos) {
    const int nb = gridDim.x, bid = blockIdx.x;
    const int xcd = bid & 7, q8 = nb >> 3, r8 = nb & 7;
    const int L = xcd * q8 + min(xcd, r8) + (bid >> 3);
    const int R = (N + 255) / 256;  // row tiles
    int lo = 0, hi = T_ - 1;        // last t with tiles[t] R <= L
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tiles[mid] * R <= L) lo = mid; else hi = mid - 1;
    }
    const int t = lo;
    const int nct = tiles[t + 1] - tiles[t], local = L - tiles[t] * R;
    const int rt = local / nct, ct = local - rt * nct;
    const int s0 = soff[t], nt = soff[t + 1] - s0, ncol = G * nt;
    const int c0 = ct * 64, r0 = rt * 256;
    const bool half = c0 + 32 >= ncol;  // uniform: only the first 32 columns carry work
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int NB16 = Np / 16, NB32 = Np / 32;

    // table fragments of the wave's two 16-row blocks (clamped: rows >= N are zero
    // fragments or, past Np, a repeated valid block whose results are not stored)
    const h8v *tp[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int ib = min((r0 + 32 * wave) / 16 + a, NB16 - 1);
        tp[a] = TABH + ((int64_t)t * NB16 + ib) * NB32 * 256 + lane;
    }
    // source tile of a K step in LDS, double-buffered, in fragment order
    // xs[stage][b * 4 + c][lane] (column block b, component c): wave w copies
    // component w of the four column blocks (columns clamped to ncol - 1)
    __shared__ h8v xs[2][16][64];
    // wave w copies component w % 4 of column blocks 2 (w / 4) and 2 (w / 4) + 1
    const _Float16 *sp[2];
    const int xc = wave & 3, xb0 = 2 * (wave >> 2);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int col = min(c0 + 16 * (xb0 + u) + (lane & 15), ncol - 1);
        const int g = col / nt, r = col - g * nt;
        sp[u] = XH + (((int64_t)g * H + smode[s0 + r]) * 4 + xc) * Np + 8 * (lane >> 4);
    }
    // the store's shifts, loaded before the K loop (in the store they would be a chain
    // of dependent loads after it): column shifts of the lane's 16 output columns
    // (clamped), row exponents of its two rows
    int csh[4][4], rsh[2];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int col = min(c0 + 16 * b + 4 * (lane >> 4) + v, ncol - 1);
            const int g = col / nt, r = col - g * nt;
            csh[b][v] = cshift[g * H + smode[s0 + r]];
        }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int i = min(r0 + 32 * wave + 16 * a + (lane & 15), N - 1);
        rsh[a] = shift_of(E[(int64_t)t * Np + i]);
    }
    f32x4v acc[3][2][4];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[p][a][b] = f32x4v{0.f, 0.f, 0.f, 0.f};

    // table fragments run PD K steps ahead in PD + 1 register sets: one step with the
    // lo halves (the register budget of two waves per SIMD), two without them (a step
    // then has a third of the MFMAs to cover the load latency with)
    constexpr int PD = LO ? 1 : 2;
    h8v tf[PD + 1][2][4], xr[2];
    // without the lo halves the waves of components 1 and 3 copy nothing (wave-uniform)
    const bool xcopy = LO || (xc & 1) == 0;
    if (xcopy) {
#pragma unroll
        for (int u = 0; u < 2; ++u) xs[0][(xb0 + u) * 4 + xc][lane] = *reinterpret_cast<const h8v *>(sp[u]);
    }
#pragma unroll
    for (int p = 0; p < PD; ++p) {
        const int jp = min(p, NB32 - 1);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int c = 0; c < 4; c += LO ? 1 : 2) tf[p][a][c] = tp[a][(int64_t)jp * 256 + c * 64];
    }
    // wait for the first steps' loads here: left pending into the loop, they would make
    // every step wait for all its own prefetches before its first MFMA (vmcnt counts
    // loads in issue order)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    // one K step on the table fragments `cur` while step jb + PD's go to `nxt` (the
    // register sets rotate, so no copy waits for the prefetch before the barrier)
    auto step = [&](int jb, h8v (&cur)[2][4], h8v (&nxt)[2][4]) {
        const int st = jb & 1;
        // the later steps' loads, issued on every step (the last ones reload the last
        // block, unused) so that every path has the same loads in flight and each wait
        // is exact; sources first: the LDS store at the end of the step then waits for
        // them only, not for the table prefetch behind them (loads complete in order)
        const int jn = min(jb + 1, NB32 - 1), jt = min(jb + PD, NB32 - 1);
        if (xcopy) {
#pragma unroll
            for (int u = 0; u < 2; ++u) xr[u] = *reinterpret_cast<const h8v *>(sp[u] + 32 * jn);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int c = 0; c < 4; c += LO ? 1 : 2) nxt[a][c] = tp[a][(int64_t)jt * 256 + c * 64];
        // acc0 += Xr Tr, acc1 += Xi Ti, acc2 += Xr Ti + Xi Tr (components 0, 1 = re hi,
        // lo; 2, 3 = im hi, lo), two column blocks at a time: consecutive MFMAs go to
        // four different accumulators, so none waits for the previous one's result
#pragma unroll
        for (int bp = 0; bp < 4; bp += 2) {
            if (bp == 2 && half) break;  // the tile's last 32 columns are all padding
            h8v xb[2][4];
#pragma unroll
            for (int d = 0; d < 2; ++d)
#pragma unroll
                for (int c = 0; c < 4; c += LO ? 1 : 2) xb[d][c] = xs[st][(bp + d) * 4 + c][lane];
#define TG_SPLIT_STEP(P, XC, TC)                                                         \
    _Pragma("unroll") for (int a = 0; a < 2; ++a)                                        \
        _Pragma("unroll") for (int d = 0; d < 2; ++d)                                    \
            acc[P][a][bp + d] = mfma_h(xb[d][XC], cur[a][TC], acc[P][a][bp + d]);
            TG_SPLIT_STEP(0, 0, 0)
            TG_SPLIT_STEP(1, 2, 2)
            TG_SPLIT_STEP(2, 0, 2)
            if (LO) {
                TG_SPLIT_STEP(0, 0, 1)
                TG_SPLIT_STEP(1, 2, 3)
                TG_SPLIT_STEP(2, 2, 0)
                TG_SPLIT_STEP(0, 1, 0)
                TG_SPLIT_STEP(1, 3, 2)
                TG_SPLIT_STEP(2, 0, 3)
                TG_SPLIT_STEP(2, 2, 1)
                TG_SPLIT_STEP(2, 1, 2)
                TG_SPLIT_STEP(2, 3, 0)
            } else {
                TG_SPLIT_STEP(2, 2, 0)
            }
#undef TG_SPLIT_STEP
        }
        if (xcopy) {
#pragma unroll
            for (int u = 0; u < 2; ++u) xs[st ^ 1][(xb0 + u) * 4 + xc][lane] = xr[u];
        }
        __syncthreads();
    };
    for (int jb = 0; jb < NB32; jb += PD + 1) {
        step(jb, tf[0], tf[PD]);
        if (jb + 1 < NB32) step(jb + 1, tf[1], tf[(1 + PD) % (PD + 1)]);
        if (PD == 2 && jb + 2 < NB32) step(jb + 2, tf[PD], tf[(2 + PD) % (PD + 1)]);
    }
    // D[m = column 4 (lane / 16) + v][n = row lane % 16]
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int i = r0 + 32 * wave + 16 * a + (lane & 15);
        if (i >= N) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int col = c0 + 16 * b + 4 * (lane >> 4) + v;
                if (col >= ncol) continue;
                const int g = col / nt, r = col - g * nt;
                const int sh = -(rsh[a] + csh[b][v]);
                U[((int64_t)g * S + gpos[s0 + r]) * N + i] =
                    cf(ldexpf(acc[0][a][b][v] - acc[1][a][b][v], sh), ldexpf(acc[2][a][b][v], sh));
            }
    }
}

// Split tables of the singular FP64 pair tables (built first when missing).
int ensure_split_tables(tg_cluster *c, const cd *tab64) {
    TG_TRY(tables_wait(c));
    if (c->have_tablesh) return TG_OK;
    const int N = c->N, T = c->plan.T, Np = (N + 31) & ~31;
    TG_TRY(c->b_texp.ensure(sizeof(int) * (size_t)T * Np));
    TG_TRY(c->b_tabh.ensure((size_t)T * Np * Np * 8));
    int *E = c->b_texp.as<int>();
    const int64_t ne = (int64_t)T * Np;
    k_fill_int<<<(unsigned)((ne + 255) / 256), 256, 0, c->stream>>>(E, ne, kNoExp);
    const int jchunk = 128;
    const dim3 grid((unsigned)((N + 255) / 256), (unsigned)T, (unsigned)((N + jchunk - 1) / jchunk));
    if (c->have_texpo)
        k_split_rowexp16<<<grid, 256, 0, c->stream>>>(c->b_texpo.as<int16_t>(), N, Np, jchunk, E);
    else
        k_split_rowexp<<<grid, 256, 0, c->stream>>>(tab64, N, Np, jchunk, E);
    TG_HIP(hipGetLastError());
    const int64_t threads = (int64_t)T * (Np / 16) * (Np / 32) * 64;
    k_split_tables<<<(unsigned)((threads + 255) / 256), 256, 0, c->stream>>>(tab64, E, N, Np, T,
                                                                             c->b_tabh.as<h8v>());
    TG_HIP(hipGetLastError());
    c->have_tablesh = true;
    c->b_texpo.release();
    c->have_texpo = false;
    return TG_OK;
}

// C x for `cols` columns on the split FP16 tables; U in c64, then the FP64 plan
// contraction of the FP32 form.  hi_only multiplies the hi halves alone.
int coupling_tables_split16(tg_cluster *c, const cd *tab64, const cd *x, cd *y, int cols, bool hi_only) {
    TG_TRY(ensure_split_tables(c, tab64));
    const int N = c->N, H = c->H, T = c->plan.T, S = c->plan.S, Np = (N + 31) & ~31;
    // right-hand sides per pass: the split sources and U of a pass within 512 MB
    const size_t per_col = (size_t)2 * H * 4 * Np * sizeof(_Float16) + sizeof(cf) * (size_t)2 * S * N;
    const int cmax = (int)std::min<size_t>(kGemmCols, std::max<size_t>(8, ((size_t)512 << 20) / per_col));
    const int ncmax = std::min(cols, cmax);
    const size_t xh_bytes = (((size_t)ncmax * 2 * H * 4 * Np * sizeof(_Float16)) + 255) & ~(size_t)255;
    const size_t sh_bytes = (((size_t)ncmax * 2 * H * sizeof(int)) + 255) & ~(size_t)255;
    TG_TRY(c->ws[0].ensure(xh_bytes + sh_bytes + sizeof(cf) * (size_t)ncmax * 2 * S * N));
    char *base = c->ws[0].as<char>();
    _Float16 *XH = reinterpret_cast<_Float16 *>(base);
    int *cshift = reinterpret_cast<int *>(base + xh_bytes);
    cf *U = reinterpret_cast<cf *>(base + xh_bytes + sh_bytes);
    Timer::Rec rec;
    TG_TRY(c->timer.begin("coupling_mf32", c->stream, rec));
    for (int c0 = 0; c0 < cols; c0 += cmax) {
        const int nc = std::min(cmax, cols - c0);
        const int Q = nc * 2 * H;
        k_split_sources<<<(unsigned)Q, 256, 0, c->stream>>>(x + (int64_t)c0 * c->n, c->n, c->nh, N, Np, H, XH,
                                                            cshift);
        TG_HIP(hipGetLastError());
        const int32_t *tiles = c->plan.stiles + (size_t)(nc - 1) * (T + 1);  // 64-wide column tiles
        const unsigned grid = (unsigned)c->gemm_ntiles[(size_t)nc - 1] * (unsigned)((N + 255) / 256);
        Timer::Rec grec;
        const char *gname = hi_only ? "gemm_split16_hi" : "gemm_split16";
        TG_TRY(c->timer.begin(gname, c->stream, grec));
        if (hi_only)
            k_zgemm_subsets_split16<false><<<grid, 512, 0, c->stream>>>(
                c->b_tabh.as<h8v>(), XH, c->b_texp.as<int>(), cshift, U, N, Np, H, S, 2 * nc, T, tiles, c->plan.soff,
                c->plan.smode, c->plan.gpos);
        else
            k_zgemm_subsets_split16<true><<<grid, 512, 0, c->stream>>>(
                c->b_tabh.as<h8v>(), XH, c->b_texp.as<int>(), cshift, U, N, Np, H, S, 2 * nc, T, tiles, c->plan.soff,
                c->plan.smode, c->plan.gpos);
        TG_HIP(hipGetLastError());
        TG_TRY(c->timer.end(gname, c->stream, grec));
        TG_TRY(launch_tab_contract_c64(c, U, nc, y + (int64_t)c0 * c->n));
    }
    TG_TRY(c->timer.end("coupling_mf32", c->stream, rec));
    return TG_OK;
}

}  // namespace tg
