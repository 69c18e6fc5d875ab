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
#include "tg_pairtab.h"

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

// Row exponents from an upper bound of each table row instead of its maximum (no pass
// over the entries): |h_p(k d)| decreases with d (|h_p(x)|^2 is a sum of positive
// powers of 1 / x), so over the sources j it is largest at the nearest other centre
// d_min(i); and the addition theorem (sum over q of |Y_p^q|^2 = (2p + 1) / (4 pi))
// bounds the unnormalised harmonic by |P_p^q| <= sqrt((p + q)! / (p - q)!), and its
// negative order (the reflection factor (p - q)! / (p + q)! times the conjugate) by
// the reciprocal square root.  E[t][i] is the exponent of that bound (times 1 + 1e-10
// against rounding in the entries), kNoExp where no other centre is apart from i
// (rows of zeros) and for the padding rows i >= N.  One wave per row.
__global__ __launch_bounds__(64) void k_split_rowexp_bound(int L2, double k, const double *pos, int N, int Np,
                                                            int *E) {
    // one wave per row: the lanes take every 64th source, then a wave minimum
    const int i = blockIdx.x, lane = threadIdx.x;
    const int T = (L2 + 1) * (L2 + 1);
    double dmin = INFINITY;
    if (i < N) {
        const double p0 = pos[3 * i], p1 = pos[3 * i + 1], p2 = pos[3 * i + 2];
        for (int j = lane; j < N; j += 64) {
            double d[3] = {p0 - pos[3 * j], p1 - pos[3 * j + 1], p2 - pos[3 * j + 2]};
            double r, u[3];
            if (j != i && unit_direction(d, r, u)) dmin = fmin(dmin, r);
        }
    }
    for (int o = 32; o > 0; o >>= 1) dmin = fmin(dmin, __shfl_xor(dmin, o));
    if (!(dmin < INFINITY)) {
        for (int t = lane; t < T; t += 64) E[(int64_t)t * Np + i] = kNoExp;
        return;
    }
    // every lane runs the Hankel recurrence up to the degrees of its harmonics
    const double x = k * dmin, inv = 1.0 / x;
    cd hprev(cos(x) * inv, sin(x) * inv), hcur(sin(x) * inv, -cos(x) * inv);  // h_{-1}, h_0
    int t = lane;
    for (int p = 0; p <= L2 && t < T; ++p) {
        const double hm = cabs_(hcur) * (1.0 + 1e-10);
        for (; t < T && t <= p * p + 2 * p; t += 64) {  // this lane's harmonics of degree p
            const int q = t - p * p - p, aq = q < 0 ? -q : q;
            double fr = 1.0;  // (p + |q|)! / (p - |q|)!
            for (int m = 1; m <= aq; ++m) fr *= (double)(p + m) * (double)(p - m + 1);
            E[(int64_t)t * Np + i] = mag_exp(q >= 0 ? hm * sqrt(fr) : hm / sqrt(fr));
        }
        const cd hnext = hankel_next((double)(2 * p + 1) * inv, hcur, hprev);
        hprev = hcur;
        hcur = hnext;
    }
}

// The FP64 pair tables TAB[t][j][i] (the matrix-free FP64 coupling's operands) and their
// split FP16 fragments TABH in one pass over the recurrences.  One workgroup per 16
// rows i x 16 sources j (half of each fragment's 32 sources: lanes 0..31 or 32..63 of
// every fragment, 512 contiguous bytes per component); thread (ii = tid % 16, jj =
// tid / 16) evaluates its pair one order q at a time with the operations of
// singular_column_put (the same values as the FP64 tables of k_pair_tables), but with
// h_0 .. h_L2 computed once into registers (LB2 >= L2 bounds the unrolled degree
// loop) and the degree-recurrence coefficients read from LDS.  Each value is stored in
// FP64 (16 consecutive rows per source: 256-byte runs) when write64, and split with
// its row's shift into LDS in fragment order; after each order the workgroup writes
// the order's fragments (up to 2 L2 harmonics x 4 components x 512 bytes) with 16-byte
// stores; an order with more harmonics than the staging holds (nsub, sized so that
// the dynamic LDS stays within the 64 KB a workgroup may allocate) is written in
// rounds.  Pairs with i or j beyond N give zero fragments (and no FP64 entries), self
// pairs and coincident centres zeros.  Dynamic LDS: the staging [nsub][4][256] halves,
// the coefficients [T] (a, b) and the row shifts [T][16] (int16).
template <int LB2>
__global__ __launch_bounds__(256) void k_pair_tables_split(DevPlan pl, double k, const double *pos, int N, int Np,
                                                           const int *E, cd *tab, h8v *TABH, int write64,
                                                           int nsub) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int T = pl.T, L2 = pl.L2;
    _Float16 *stage = reinterpret_cast<_Float16 *>(lds);     // [nsub][4][256]
    double2 *coef = reinterpret_cast<double2 *>(lds + (size_t)nsub * 4 * 256 * sizeof(_Float16));  // [T]
    int16_t *shift = reinterpret_cast<int16_t *>(coef + T);  // [T][16]
    const int tid = threadIdx.x, ii = tid & 15, jj = tid >> 4;
    const int j16 = blockIdx.x, ib = blockIdx.y;
    const int i = ib * 16 + ii, j = j16 * 16 + jj;
    const int NB16 = Np / 16, NB32 = Np / 32, jb = j16 >> 1, jhalf = j16 & 1;
    const int64_t NN = (int64_t)N * N;
    for (int e = tid; e < T * 16; e += 256) {
        const int t = e >> 4, row = ib * 16 + (e & 15);
        shift[e] = (int16_t)(row < N ? shift_of(E[(int64_t)t * Np + row]) : 0);
    }
    for (int h = tid; h < T; h += 256) {
        const int p = (int)sqrt((double)h + 0.5), q = h - p * p - p;
        double2 ab = make_double2(0.0, 0.0);
        if (q >= 0 && p > q) ab = make_double2((double)(2 * p - 1) / (double)(p - q),
                                               (double)(p + q - 1) / (double)(p - q));
        coef[h] = ab;
    }
    // this thread's place in the fragments: lane (ii + 16 (jj / 8)) of the half, element
    // jj % 8
    const int slot = (ii + 16 * (jj >> 3)) * 8 + (jj & 7);
    const bool valid = i < N && j < N;  // an entry of the FP64 tables
    double d[3], r = 1.0, u[3] = {0.0, 0.0, 1.0};
    bool live = valid && i != j;
    if (live) {
        for (int a = 0; a < 3; ++a) d[a] = pos[3 * i + a] - pos[3 * j + a];
        live = unit_direction(d, r, u);
    }
    cd hv[LB2 + 1];
    cd xy(u[0], u[1]), ladder(1.0);
    {
        const double x = k * r, inv = 1.0 / x, sx = sin(x), cx_ = cos(x);
        cd hprev(cx_ * inv, sx * inv), hcur(sx * inv, -cx_ * inv);  // h_{-1}, h_0
#pragma clang loop unroll(full)
        for (int p = 0; p <= LB2; ++p) {
            hv[p] = hcur;
            const cd hnext = hankel_next((double)(2 * p + 1) * inv, hcur, hprev);
            hprev = hcur;
            hcur = hnext;
        }
    }
    cd *tij = tab + (int64_t)j * N + i;
    // write the staged fragments of the order's staging indices base .. base + count - 1
    // (held at staging slots 0 .. count - 1): chunk = (slot, component), 32 lanes of 16
    // bytes each.  Every thread calls it at the same points (the put sequence of an
    // order is the same for all), so its barriers are uniform.
    auto flush = [&](int q, int base, int count) {
        __syncthreads();
        const h8v *sv = reinterpret_cast<const h8v *>(stage);
        for (int w = tid; w < count * 4 * 32; w += 256) {
            const int ch = w >> 5, l32 = w & 31, sidx = base + (ch >> 2), comp = ch & 3;
            const int p = q + (q ? sidx >> 1 : sidx);
            const int h = p * p + p + ((q && (sidx & 1)) ? -q : q);
            TABH[(((int64_t)h * NB16 + ib) * NB32 + jb) * 256 + comp * 64 + jhalf * 32 + l32] = sv[ch * 32 + l32];
        }
        __syncthreads();
    };
    __syncthreads();
    for (int q = 0; q <= L2; ++q) {
        // (p, +q) and then (p, -q) for p = q, q + 1, ...: the staging index is the
        // call count; the staging holds nsub of them at a time
        int s = 0, base = 0;
        auto put = [&](int h, cd v) {
            if (write64 && valid) {
                typedef double d2v __attribute__((ext_vector_type(2)));
                const d2v x = {v.re, v.im};
                __builtin_nontemporal_store(x, reinterpret_cast<d2v *>(tij + (int64_t)h * NN));
            }
            const int sh = shift[h * 16 + ii];
            const half2s pr = split16(v.re, sh), pi = split16(v.im, sh);
            _Float16 *o = stage + (size_t)(s - base) * 1024 + slot;
            o[0] = pr.hi;
            o[256] = pr.lo;
            o[512] = pi.hi;
            o[768] = pi.lo;
            if (++s - base == nsub) {
                flush(q, base, nsub);
                base = s;
            }
        };
        if (q > 0) ladder = (-(double)(2 * q - 1)) * cmul_fma(ladder, xy);
        cd cur = ladder, prev(0.0);
        // fully unrolled, so hv[p] is a register, never scratch
#pragma clang loop unroll(full)
        for (int p = 0; p <= LB2; ++p) {
            if (p > L2 || p < q) continue;
            if (p > q) {
                const double2 ab = coef[p * p + p + q];
                const cd next = legendre_next(ab.x * u[2], ab.y, cur, prev);
                prev = cur;
                cur = next;
            }
            // values chosen per lane, put called by every thread alike: put may flush,
            // and its barriers must not sit in a divergent branch
            const cd vpos = live ? cmul_fma(hv[p], cur) : cd(0.0);
            put(p * p + p + q, vpos);
            if (q > 0) {
                const cd vneg = live ? cmul_fma(hv[p], reflect(pl.refl[p * p + p - q], cur)) : cd(0.0);
                put(p * p + p - q, vneg);
            }
        }
        if (s > base) flush(q, base, s - base);
    }
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

// Split tables (and the FP64 tables when they are missing) in one pass over the
// recurrences, with row exponents from the bound of k_split_rowexp_bound.
int ensure_split_tables(tg_cluster *c) {
    TG_TRY(split_wait(c));
    if (c->have_tablesh) return TG_OK;
    const int N = c->N, T = c->plan.T, L2 = c->plan.L2, Np = (N + 31) & ~31;
    const bool write64 = !c->have_tables;
    if (write64) {
        TG_TRY(tables_wait(c));
        TG_TRY(c->b_tab.ensure(sizeof(cd) * (size_t)N * N * T));
    }
    TG_TRY(c->b_texp.ensure(sizeof(int) * (size_t)T * Np));
    TG_TRY(c->b_tabh.ensure((size_t)T * Np * Np * 8));
    int *E = c->b_texp.as<int>();
    // staging for up to the most harmonics of one order, within 64 KB of dynamic LDS
    // beside the coefficients and the row shifts
    const size_t fixed = sizeof(double2) * (size_t)T + sizeof(int16_t) * (size_t)T * 16;
    const size_t per = sizeof(_Float16) * 4 * 256;
    const int nstage = std::max(2 * L2, L2 + 1);
    const int nsub = (int)std::min<size_t>(nstage, fixed < 65536 ? (65536 - fixed) / per : 0);
    if (nsub < 1) {
        set_error("split tables: lmax %d needs more LDS than a workgroup may allocate", c->plan.lmax);
        return TG_INVALID_INPUT;
    }
    const size_t lds = per * (size_t)nsub + fixed;
    k_split_rowexp_bound<<<(unsigned)Np, 64, 0, c->stream>>>(L2, c->k0, c->b_pos.as<double>(), N, Np, E);
    TG_HIP(hipGetLastError());
    const dim3 grid((unsigned)(Np / 16), (unsigned)(Np / 16));
#define TG_TABLES_SPLIT(LB)                                                                                     \
    k_pair_tables_split<LB><<<grid, 256, lds, c->stream>>>(c->plan, c->k0, c->b_pos.as<double>(), N, Np, E,    \
                                                          c->b_tab.as<cd>(), c->b_tabh.as<h8v>(), write64 ? 1 : 0, nsub)
    if (L2 <= 6) TG_TABLES_SPLIT(6);
    else if (L2 <= 12) TG_TABLES_SPLIT(12);
    else TG_TABLES_SPLIT(24);
#undef TG_TABLES_SPLIT
    TG_HIP(hipGetLastError());
    c->have_tablesh = true;
    if (write64) c->have_tables = true;
    return TG_OK;
}

// C x for `cols` columns on the split FP16 tables; U in c64, then the FP64 plan
// contraction of the FP32 form.  hi_only multiplies the hi halves alone.
int coupling_tables_split16(tg_cluster *c, const cd *x, cd *y, int cols, bool hi_only) {
    TG_TRY(ensure_split_tables(c));
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
