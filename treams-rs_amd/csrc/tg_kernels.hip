// gfx950 kernels of the cluster-coupling path.
//
// Layouts.  A vector of n = N M unknowns is kept in the helicity-split layout
// [h][N*H] with h = 0 for pol 1 (positive helicity) and h = 1 for pol 0, H = M / 2
// modes (l, m) per sphere and helicity, so the same-helicity coupling is two
// N H x N H matrices C_+ and C_-.  Stored couplings are column-major, source
// column (j H + a), destination row (i H + b), with (l, m) index a = l^2 + l + m - 1.
#include <algorithm>
#include "tg_internal.h"
#include "tg_mfma.h"
#include "tg_pairtab.h"
#include "tg_special.h"

namespace tg {

// ---------------------------------------------------------------- geometry --

// Radial values z_p(kr), p = 0..L2: outgoing Hankel or regular Bessel.
__device__ void radial_values(int L2, double kr, bool regular, cd *rad) {
    if (!regular) {
        tg_hankel_seq(L2, kr, rad);
    } else {
        double j[64];
        tg_bessel_j_seq(L2, kr, j);
        for (int p = 0; p <= L2; ++p) rad[p] = cd(j[p]);
    }
}

// Entries (p, +-q), p = q..L2, of the harmonic table z_p(kr) P_p^q e^{i q phi} at the
// unit vector u: diagonal ladder P_q^q = (-1)^q (2q-1)!! (x + i y)^q, then the degree
// recurrence (p - q) S_p = (2p - 1) z S_{p-1} - (p + q - 1) S_{p-2}; negative orders by
// S_p^{-q} = (-1)^q (p-q)!/(p+q)! conj(S_p^q).
// put(h, value) stores harmonic index h.
template <class Put>
__device__ __forceinline__ void table_column_put(int q, int L2, const double u[3], const cd *rad, const double *refl,
                                                 Put put) {
    cd xy(u[0], u[1]);
    cd cur(1.0);
    for (int k = 1; k <= q; ++k) cur = (-(double)(2 * k - 1)) * (cur * xy);
    cd prev(0.0);
    for (int p = q; p <= L2; ++p) {
        if (p > q) {
            double a = (double)(2 * p - 1) / (double)(p - q);
            double b = (double)(p + q - 1) / (double)(p - q);
            cd next = (a * u[2]) * cur - b * prev;
            prev = cur;
            cur = next;
        }
        put(p * p + p + q, rad[p] * cur);
        if (q > 0) put(p * p + p - q, rad[p] * (refl[p * p + p - q] * conj(cur)));
    }
}

__device__ void table_column(int q, int L2, const double u[3], const cd *rad, const double *refl, cd *tab,
                             int64_t stride = 1) {
    table_column_put(q, L2, u, rad, refl, [&](int h, cd v) { tab[(int64_t)h * stride] = v; });
}

template <class Put>
__device__ __forceinline__ void zero_column_put(int q, int L2, Put put) {
    for (int p = q; p <= L2; ++p) {
        put(p * p + p + q, cd(0.0));
        if (q > 0) put(p * p + p - q, cd(0.0));
    }
}

__device__ __forceinline__ void zero_column(int q, int L2, cd *tab, int64_t stride = 1) {
    zero_column_put(q, L2, [&](int h, cd v) { tab[(int64_t)h * stride] = v; });
}

// Both helicity values of one block entry from the plan and a harmonic table.
__device__ __forceinline__ void entry_values(const DevPlan &pl, int e, const cd *tab, cd &cp, cd &cm) {
    const int s0 = pl.starts[e], s1 = pl.starts[e + 1], ns = pl.nsame[e];
    cd same(0.0), cross(0.0);
    for (int t = s0; t < s0 + ns; ++t) same += pl.w[t] * tab[pl.hidx[t]];
    for (int t = s0 + ns; t < s1; ++t) cross += pl.w[t] * tab[pl.hidx[t]];
    cp = cd(same.re - cross.im, same.im + cross.re);  // same + i cross
    cm = cd(same.re + cross.im, same.im - cross.re);  // same - i cross
}

// ------------------------------------------------------------------- Mie --

// One thread per sphere: 2x2 helicity Mie blocks (mode order, row-major) of every
// degree and their derivatives along the size parameter and the permittivity.
// One thread per (sphere, degree): the helicity blocks of the Mie coefficients and
// their size and permittivity derivatives (NaN for a zero denominator).
__global__ void k_mie(int N, int lmax, double k0, const double *radii, const double *eps, cd *val, cd *dsz, cd *deps) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * lmax) return;
    const int i = t / lmax, l = t - i * lmax;
    tg_dual a, b;
    const int rc = tg_mie_bh(l + 1, lmax, k0 * radii[i], cd(eps[2 * i], eps[2 * i + 1]), a, b);
    cd d = -0.5 * (a.v + b.v), o = -0.5 * (a.v - b.v);
    cd ddx = -0.5 * (a.dx + b.dx), odx = -0.5 * (a.dx - b.dx);
    cd dde = -0.5 * (a.de + b.de), ode = -0.5 * (a.de - b.de);
    if (rc) d = o = ddx = odx = dde = ode = cd(NAN, NAN);
    cd *v = val + (size_t)t * 4, *x = dsz + (size_t)t * 4, *e = deps + (size_t)t * 4;
    v[0] = d; v[1] = o; v[2] = o; v[3] = d;
    x[0] = ddx; x[1] = odx; x[2] = odx; x[3] = ddx;
    e[0] = dde; e[1] = ode; e[2] = ode; e[3] = dde;
}

int launch_mie(tg_cluster *c) {
    const double *pos = c->b_pos.as<double>();
    const double *radii = pos + 3 * c->N;
    const double *eps = radii + c->N;
    cd *val = c->b_mie.as<cd>();
    size_t blk = (size_t)c->N * c->lmax * 4;
    k_mie<<<(c->N * c->lmax + 63) / 64, 64, 0, c->stream>>>(c->N, c->lmax, c->k0, radii, eps, val, val + blk,
                                                             val + 2 * blk);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// -------------------------------------------------------------- assembly --

struct GeoCluster {
    const double *pos;
    int N;
    // 0: no row, 1: coincident (self), 2: regular pair with displacement d
    __device__ int get(int col, int row, double d[3]) const {
        if (row >= N) return 0;
        if (row == col) return 1;
        for (int a = 0; a < 3; ++a) d[a] = pos[3 * row + a] - pos[3 * col + a];
        return 2;
    }
};

struct GeoList {
    const double *disp;
    __device__ int get(int col, int row, double d[3]) const {
        if (row != 0) return 0;
        for (int a = 0; a < 3; ++a) d[a] = disp[3 * (size_t)col + a];
        return 2;
    }
};

// A displacement list taken P at a time: workgroup `col` holds displacements col P ..
// col P + P - 1 (its rows), none past `count`.
template <int P>
struct GeoListP {
    const double *disp;
    int64_t count;
    __device__ int get(int col, int row, double d[3]) const {
        const int64_t idx = (int64_t)col * P + row;
        if (idx >= count) return 0;
        for (int a = 0; a < 3; ++a) d[a] = disp[3 * idx + a];
        return 2;
    }
};

// Streaming (nontemporal) stores for data written once and read much later: the
// stored coupling, the operator of the LU and the pair tables are all far larger than
// the caches.
typedef double tg_d2v __attribute__((ext_vector_type(2)));
typedef float tg_f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void stream_store(double2 *p, double2 v) {
    tg_d2v x = {v.x, v.y};
    __builtin_nontemporal_store(x, reinterpret_cast<tg_d2v *>(p));
}
__device__ __forceinline__ void stream_store(float2 *p, float2 v) {
    tg_f2v x = {v.x, v.y};
    __builtin_nontemporal_store(x, reinterpret_cast<tg_f2v *>(p));
}
__device__ __forceinline__ void stream_store(cd *p, cd v) {
    tg_d2v x = {v.re, v.im};
    __builtin_nontemporal_store(x, reinterpret_cast<tg_d2v *>(p));
}
__device__ __forceinline__ void stream_store(cf *p, cf v) {
    tg_f2v x = {v.re, v.im};
    __builtin_nontemporal_store(x, reinterpret_cast<tg_f2v *>(p));
}

template <class S> __device__ __forceinline__ S cvt(cd v);
template <> __device__ __forceinline__ float2 cvt<float2>(cd v) { return make_float2((float)v.re, (float)v.im); }
template <> __device__ __forceinline__ double2 cvt<double2>(cd v) { return make_double2(v.re, v.im); }

template <class S>
struct OutSplit {
    S *cp, *cm;
    int64_t nh;
    int H;
    __device__ void store(int i, int j, int a, int b, cd vp, cd vm) const {
        int64_t idx = ((int64_t)j * H + a) * nh + (int64_t)i * H + b;
        stream_store(cp + idx, cvt<S>(vp));
        stream_store(cm + idx, cvt<S>(vm));
    }
};

// Reference layout: rows i M + 2 b + (1 - pol), columns j M + 2 a + (1 - pol);
// opposite-helicity entries stay at the caller's zero fill.
struct OutRef {
    double2 *out;
    int64_t n;
    int M;
    __device__ void store(int i, int j, int a, int b, cd vp, cd vm) const {
        int64_t row = (int64_t)i * M + 2 * b, col = (int64_t)j * M + 2 * a;
        out[col * n + row] = make_double2(vp.re, vp.im);
        out[(col + 1) * n + row + 1] = make_double2(vm.re, vm.im);
    }
};

// The operator A = I - T C itself, rows and columns in the split order
// (h N H + i H + b), column-major with leading dimension n:
// A[(h, i, b), (h', j, a)] = delta - T_i[h, h'](l_b) C_h'[(i, b), (j, a)].
template <class S>
struct OutOperator {
    S *A;
    int64_t n, nh;
    int H, lmax;
    const cd *mie;
    __device__ void store(int i, int j, int a, int b, cd vp, cd vm) const {
        int l = 1;
        while ((l + 1) * (l + 1) <= b + 1) ++l;  // b = l^2 + l + m - 1
        const cd *t = mie + (size_t)(i * lmax + l - 1) * 4;
        const int64_t r = (int64_t)i * H + b, col = (int64_t)j * H + a;
        const double d = (i == j && a == b) ? 1.0 : 0.0;
        stream_store(A + col * n + r, cvt<S>(cd(d) - t[0] * vp));
        stream_store(A + (col + nh) * n + r, cvt<S>(-(t[1] * vm)));
        stream_store(A + col * n + r + nh, cvt<S>(-(t[2] * vp)));
        stream_store(A + (col + nh) * n + r + nh, cvt<S>(cd(d) - t[3] * vm));
    }
};

// One M x M block per list entry.
struct OutBlocks {
    double2 *out;
    int M;
    __device__ void store(int, int j, int a, int b, cd vp, cd vm) const {
        double2 *blk = out + (size_t)j * M * M;
        blk[(size_t)(2 * a) * M + 2 * b] = make_double2(vp.re, vp.im);
        blk[(size_t)(2 * a + 1) * M + 2 * b + 1] = make_double2(vm.re, vm.im);
    }
};

// The two same-helicity H x H blocks of each pair, [pair][h][a][b] (destination b inner;
// h = 0 the reference's first polarisation index; the opposite-helicity entries are zero
// and not stored), each complex value as two bfloat16 (real part in the low half, rounded
// to nearest even): the near-field blocks of the preconditioner applied inside
// mixed-precision solves, where an approximate operator only costs iterations.
__device__ __forceinline__ uint32_t bf16_bits(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;  // finite values
}
struct OutBlocksHelicityBF16 {
    uint32_t *out;
    int H;
    __device__ void store(int, int j, int a, int b, cd vp, cd vm) const {
        uint32_t *blk = out + (size_t)j * 2 * H * H;
        blk[(size_t)a * H + b] = bf16_bits((float)vp.re) | (bf16_bits((float)vp.im) << 16);
        blk[(size_t)H * H + (size_t)a * H + b] = bf16_bits((float)vm.re) | (bf16_bits((float)vm.im) << 16);
    }
};

// The same for GeoListP: pair col P + row.
template <int P>
struct OutBlocksHelicityBF16P {
    uint32_t *out;
    int H;
    __device__ void store(int row, int col, int a, int b, cd vp, cd vm) const {
        OutBlocksHelicityBF16{out, H}.store(0, col * P + row, a, b, vp, vm);
    }
};

// Assembly: one workgroup per (source unit, group of P destination units).  The
// harmonic tables of the P pairs sit in LDS; each thread owns block entries and
// reuses every plan term it loads across the P pairs, accumulating in FP64.  The
// radial and Legendre recurrences run in FP64; with TabT = float2 the finished tables
// are staged in LDS rounded to FP32 (for the c64 stored coupling, whose entries are
// rounded to FP32 anyway), which halves the LDS traffic that bounds the contraction.
template <class TabT> __device__ __forceinline__ cd tab_get(const TabT *t, int i);
template <> __device__ __forceinline__ cd tab_get<double2>(const double2 *t, int i) {
    const double2 v = t[i];
    return cd(v.x, v.y);
}
template <> __device__ __forceinline__ cd tab_get<float2>(const float2 *t, int i) {
    const float2 v = t[i];
    return cd((double)v.x, (double)v.y);
}

template <int P, class Geo, class Out, bool REG, class TabT = double2>
__global__ __launch_bounds__(256) void k_assemble(DevPlan pl, double k, Geo geo, Out out) {
    extern __shared__ cd smem[];
    const int T = pl.T, L2 = pl.L2, H = pl.H;
    constexpr bool copy = sizeof(TabT) != sizeof(cd);
    cd *tab = smem;                 // FP64 tables [P][T] while they are built
    cd *rad = smem + P * T;
    // [P][T] tables read by the contraction: the FP64 tables themselves, or an FP32 copy
    TabT *tabs = copy ? reinterpret_cast<TabT *>(rad + P * (L2 + 1)) : reinterpret_cast<TabT *>(tab);
    __shared__ double su[P][3];
    __shared__ int flag[P];
    const int col = blockIdx.x;
    const int row0 = blockIdx.y * P;
    if (threadIdx.x < P) {
        const int il = threadIdx.x;
        double d[3], r, u[3];
        int f = geo.get(col, row0 + il, d);
        if (f == 2 && !unit_direction(d, r, u)) f = 1;
        if (f == 2) {
            radial_values(L2, k * r, REG, rad + il * (L2 + 1));
            su[il][0] = u[0]; su[il][1] = u[1]; su[il][2] = u[2];
        }
        flag[il] = f;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < P * (L2 + 1); t += blockDim.x) {
        const int il = t / (L2 + 1), q = t - il * (L2 + 1);
        cd *tb = tab + il * T;
        if (flag[il] == 2) {
            double u[3] = {su[il][0], su[il][1], su[il][2]};
            table_column(q, L2, u, rad + il * (L2 + 1), pl.refl, tb);
        } else {
            zero_column(q, L2, tb);
            if (REG && q == 0 && flag[il] == 1) tb[0] = cd(1.0);  // regular plan at the origin
        }
    }
    __syncthreads();
    if (copy) {
        for (int t = threadIdx.x; t < P * T; t += blockDim.x) {
            TabT v;
            v.x = tab[t].re;
            v.y = tab[t].im;
            tabs[t] = v;
        }
        __syncthreads();
    }
    const int HH = H * H;
    for (int e = threadIdx.x; e < HH; e += blockDim.x) {
        const int a = e / H, b = e - a * H;
        const int s0 = pl.starts[e], s1 = pl.starts[e + 1], ns = pl.nsame[e];
        cd same[P], cross[P];
#pragma unroll
        for (int il = 0; il < P; ++il) { same[il] = cd(0.0); cross[il] = cd(0.0); }
        for (int t = s0; t < s0 + ns; ++t) {
            const double wt = pl.w[t];
            const int h = pl.hidx[t];
#pragma unroll
            for (int il = 0; il < P; ++il) same[il] += wt * tab_get(tabs, il * T + h);
        }
        for (int t = s0 + ns; t < s1; ++t) {
            const double wt = pl.w[t];
            const int h = pl.hidx[t];
#pragma unroll
            for (int il = 0; il < P; ++il) cross[il] += wt * tab_get(tabs, il * T + h);
        }
#pragma unroll
        for (int il = 0; il < P; ++il) {
            if (flag[il] == 0) continue;
            cd vp(same[il].re - cross[il].im, same[il].im + cross[il].re);
            cd vm(same[il].re + cross[il].im, same[il].im - cross[il].re);
            out.store(row0 + il, col, a, b, vp, vm);
        }
    }
}

// The dense c64 store, a write stream of 2 (N H)^2 FP32 pairs: one workgroup per
// (source sphere, P destination spheres).  The harmonic tables go straight from the
// FP64 recurrences into FP32 LDS and the terms accumulate in FP32 (the stored entries
// are FP32): half the LDS of FP64 tables and half the registers of FP64
// accumulators, so six waves per SIMD (P = 16) keep the stores going while others
// compute.  Each plan term a thread loads serves its P destination spheres.
template <int P>
__global__ __launch_bounds__(256) void k_assemble_c64(DevPlan pl, double k, GeoCluster geo, float2 *cp, float2 *cm,
                                                      int64_t nh) {
    extern __shared__ float2 smf[];
    const int T = pl.T, L2 = pl.L2, H = pl.H;
    float2 *tabs = smf;                                       // [P][T]
    cd *rad = reinterpret_cast<cd *>(smf + ((P * T + 1) & ~1));  // [P][L2 + 1]
    __shared__ double su[P][3];
    __shared__ int flag[P];
    const int col = blockIdx.x;
    const int row0 = blockIdx.y * P;
    if (threadIdx.x < P) {
        const int il = threadIdx.x;
        double d[3], r, u[3];
        int f = geo.get(col, row0 + il, d);
        if (f == 2 && !unit_direction(d, r, u)) f = 1;
        if (f == 2) {
            radial_values(L2, k * r, false, rad + il * (L2 + 1));
            su[il][0] = u[0]; su[il][1] = u[1]; su[il][2] = u[2];
        }
        flag[il] = f;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < P * (L2 + 1); t += blockDim.x) {
        const int il = t / (L2 + 1), q = t - il * (L2 + 1);
        float2 *tb = tabs + il * T;
        if (flag[il] == 2) {
            double u[3] = {su[il][0], su[il][1], su[il][2]};
            table_column_put(q, L2, u, rad + il * (L2 + 1), pl.refl,
                             [&](int h, cd v) { tb[h] = make_float2((float)v.re, (float)v.im); });
        } else {
            zero_column_put(q, L2, [&](int h, cd) { tb[h] = make_float2(0.f, 0.f); });
        }
    }
    __syncthreads();
    const int HH = H * H;
    for (int e = threadIdx.x; e < HH; e += blockDim.x) {
        const int a = e / H, b = e - a * H;
        const int s0 = pl.starts[e], s1 = pl.starts[e + 1], ns = pl.nsame[e];
        float2 same[P], cross[P];
#pragma unroll
        for (int il = 0; il < P; ++il) { same[il] = make_float2(0.f, 0.f); cross[il] = make_float2(0.f, 0.f); }
        for (int t = s0; t < s0 + ns; ++t) {
            const float wt = (float)pl.w[t];
            const int h = pl.hidx[t];
#pragma unroll
            for (int il = 0; il < P; ++il) {
                const float2 v = tabs[il * T + h];
                same[il].x = fmaf(wt, v.x, same[il].x);
                same[il].y = fmaf(wt, v.y, same[il].y);
            }
        }
        for (int t = s0 + ns; t < s1; ++t) {
            const float wt = (float)pl.w[t];
            const int h = pl.hidx[t];
#pragma unroll
            for (int il = 0; il < P; ++il) {
                const float2 v = tabs[il * T + h];
                cross[il].x = fmaf(wt, v.x, cross[il].x);
                cross[il].y = fmaf(wt, v.y, cross[il].y);
            }
        }
        const int64_t base = ((int64_t)col * H + a) * nh + (int64_t)row0 * H + b;
#pragma unroll
        for (int il = 0; il < P; ++il) {
            if (flag[il] == 0) continue;
            tg_f2v vp = {same[il].x - cross[il].y, same[il].y + cross[il].x};
            tg_f2v vm = {same[il].x + cross[il].y, same[il].y - cross[il].x};
            __builtin_nontemporal_store(vp, reinterpret_cast<tg_f2v *>(cp + base + il * H));
            __builtin_nontemporal_store(vm, reinterpret_cast<tg_f2v *>(cm + base + il * H));
        }
    }
}

// FP64 tables and radial values, then the contraction copy of the tables (TabT)
template <class TabT = double2>
static size_t assemble_smem(const DevPlan &pl, int P) {
    const size_t copy = sizeof(TabT) != sizeof(cd) ? (size_t)P * pl.T * sizeof(TabT) : 0;
    return (size_t)P * (pl.T + pl.L2 + 1) * sizeof(cd) + copy;
}

// Dynamic LDS of one assembly workgroup must fit the 64 KB a workgroup may allocate.
static int check_smem(size_t bytes, int lmax) {
    if (bytes > 65536) {
        set_error("assembly at lmax %d needs %zu bytes of LDS per workgroup (limit 65536)", lmax, bytes);
        return TG_INVALID_INPUT;
    }
    return TG_OK;
}

int launch_assemble_split(tg_cluster *c, int storage, void *outp) {
    constexpr int P = 8;
    dim3 grid(c->N, (c->N + P - 1) / P);
    GeoCluster geo{c->b_pos.as<double>(), c->N};
    size_t smem = assemble_smem(c->plan, P);
    Timer::Rec rec;
    TG_TRY(c->timer.begin("assemble", c->stream, rec));
    TG_TRY(check_smem(smem, c->lmax));
    // the c64 kernel: 16 destination spheres per workgroup, measured 7.3 ms at N = 1024,
    // lmax 6 against 7.7 ms for 8 and 7.4 ms for 24 or 32 (fewer waves per SIMD)
    constexpr int PC = 16;
    const size_t smemf = (size_t)((PC * c->plan.T + 1) & ~1) * sizeof(float2) +
                         (size_t)PC * (c->plan.L2 + 1) * sizeof(cd);
    if (storage == TG_STORE_C64 && (smemf > 65536 || !fp32_tables_in_range(c))) {
        // high degrees, or tables beyond FP32 range (close spheres): FP64 tables in LDS
        float2 *base = (float2 *)outp;
        OutSplit<float2> out{base, base + c->nh * c->nh, c->nh, c->H};
        k_assemble<P, GeoCluster, OutSplit<float2>, false><<<grid, 256, smem, c->stream>>>(c->plan, c->k0, geo, out);
    } else if (storage == TG_STORE_C64) {
        float2 *base = (float2 *)outp;
        k_assemble_c64<PC><<<dim3(c->N, (c->N + PC - 1) / PC), 256, smemf, c->stream>>>(c->plan, c->k0, geo, base,
                                                                                        base + c->nh * c->nh, c->nh);
    } else {
        double2 *base = (double2 *)outp;
        OutSplit<double2> out{base, base + c->nh * c->nh, c->nh, c->H};
        k_assemble<P, GeoCluster, OutSplit<double2>, false><<<grid, 256, smem, c->stream>>>(c->plan, c->k0, geo, out);
    }
    TG_HIP(hipGetLastError());
    TG_TRY(c->timer.end("assemble", c->stream, rec));
    return TG_OK;
}

int launch_assemble_operator(tg_cluster *c, int precision, void *outp) {
    if (c->particles) return assemble_operator_particles(c, precision, outp);
    constexpr int P = 8;
    dim3 grid(c->N, (c->N + P - 1) / P);
    GeoCluster geo{c->b_pos.as<double>(), c->N};
    size_t smem = assemble_smem(c->plan, P);
    TG_TRY(check_smem(smem, c->lmax));
    Timer::Rec rec;
    TG_TRY(c->timer.begin("assemble_operator", c->stream, rec));
    if (precision == TG_STORE_C64) {
        OutOperator<float2> out{(float2 *)outp, c->n, c->nh, c->H, c->lmax, c->b_mie.as<cd>()};
        const size_t smem32 = assemble_smem<float2>(c->plan, P);
        if (smem32 <= 65536 && fp32_tables_in_range(c))
            k_assemble<P, GeoCluster, OutOperator<float2>, false, float2>
                <<<grid, 256, smem32, c->stream>>>(c->plan, c->k0, geo, out);
        else
            k_assemble<P, GeoCluster, OutOperator<float2>, false><<<grid, 256, smem, c->stream>>>(c->plan, c->k0,
                                                                                                  geo, out);
    } else {
        OutOperator<double2> out{(double2 *)outp, c->n, c->nh, c->H, c->lmax, c->b_mie.as<cd>()};
        k_assemble<P, GeoCluster, OutOperator<double2>, false><<<grid, 256, smem, c->stream>>>(c->plan, c->k0, geo,
                                                                                               out);
    }
    TG_HIP(hipGetLastError());
    TG_TRY(c->timer.end("assemble_operator", c->stream, rec));
    return TG_OK;
}

int launch_assemble_ref(tg_cluster *c, int regular, cd *outp) {
    constexpr int P = 4;
    TG_HIP(hipMemsetAsync(outp, 0, sizeof(cd) * c->n * c->n, c->stream));
    dim3 grid(c->N, (c->N + P - 1) / P);
    GeoCluster geo{c->b_pos.as<double>(), c->N};
    OutRef out{(double2 *)outp, c->n, c->M};
    size_t smem = assemble_smem(c->plan, P);
    TG_TRY(check_smem(smem, c->lmax));
    if (regular)
        k_assemble<P, GeoCluster, OutRef, true><<<grid, 256, smem, c->stream>>>(c->plan, c->k0, geo, out);
    else
        k_assemble<P, GeoCluster, OutRef, false><<<grid, 256, smem, c->stream>>>(c->plan, c->k0, geo, out);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// Zero `count` doubles with vector stores (a kernel, so code that a caller captures
// into a graph records no memset node).
__global__ void k_zero_f64(double *p, int64_t count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = 0.0;
}

int launch_zero_f64(double *p, int64_t count, hipStream_t s) {
    if (count <= 0) return TG_OK;
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 4096);
    k_zero_f64<<<(unsigned)blocks, 256, 0, s>>>(p, count);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

int launch_translation_blocks(int lmax, const DevPlan &pl, double k, int64_t count, const double *disp, int regular,
                              cd *outp, hipStream_t s) {
    const int M = 2 * lmax * (lmax + 2);
    TG_TRY(launch_zero_f64((double *)outp, (int64_t)2 * M * M * count, s));
    if (count == 0) return TG_OK;
    dim3 grid((unsigned)count, 1);
    GeoList geo{disp};
    OutBlocks out{(double2 *)outp, M};
    size_t smem = assemble_smem(pl, 1);
    TG_TRY(check_smem(smem, lmax));
    if (regular)
        k_assemble<1, GeoList, OutBlocks, true><<<grid, 256, smem, s>>>(pl, k, geo, out);
    else
        k_assemble<1, GeoList, OutBlocks, false><<<grid, 256, smem, s>>>(pl, k, geo, out);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

int launch_translation_blocks_helicity_bf16(int lmax, const DevPlan &pl, double k, int64_t count,
                                           const double *disp, uint32_t *outp, hipStream_t s, bool fp32_ok) {
    // 8 pairs per workgroup (each plan term loaded once for the 8), their tables in
    // LDS rounded to FP32 (the blocks are stored as bfloat16); every entry of every
    // block is written, so the output needs no clearing
    constexpr int P = 8;
    const int H = lmax * (lmax + 2);
    if (count == 0) return TG_OK;
    const size_t smem = assemble_smem<float2>(pl, P);
    if (smem <= 65536 && fp32_ok) {
        GeoListP<P> geo{disp, count};
        OutBlocksHelicityBF16P<P> out{outp, H};
        k_assemble<P, GeoListP<P>, OutBlocksHelicityBF16P<P>, false, float2>
            <<<dim3((unsigned)((count + P - 1) / P), 1), 256, smem, s>>>(pl, k, geo, out);
    } else {
        // high degrees, or tables beyond FP32 range: one pair per workgroup, FP64 tables
        const size_t smem1 = assemble_smem(pl, 1);
        TG_TRY(check_smem(smem1, lmax));
        k_assemble<1, GeoList, OutBlocksHelicityBF16, false><<<dim3((unsigned)count, 1), 256, smem1, s>>>(
            pl, k, GeoList{disp}, OutBlocksHelicityBF16{outp, H});
    }
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// ----------------------------------------- translation blocks, rectangular plans --

// One workgroup per displacement: the harmonic table z_p(kd) P_p^q e^{i q phi},
// p <= L2, in LDS (regular: identity at the origin), then every (source a,
// destination b) entry of the Md x Ms block: column-major, destination inner, in the
// reference mode order; opposite helicities stay at the caller's zero fill.
template <bool REG>
__global__ __launch_bounds__(256) void k_translation_rect(RectPlan pl, double k, const double *disp, cd *out) {
    extern __shared__ cd rect_lds[];
    cd *tab = rect_lds, *rad = rect_lds + pl.T;
    __shared__ int ok_s;
    __shared__ double u_s[3];
    const int blk = blockIdx.x;
    if (threadIdx.x == 0) {
        double d[3] = {disp[3 * blk], disp[3 * blk + 1], disp[3 * blk + 2]}, r, u[3];
        ok_s = unit_direction(d, r, u);
        if (ok_s) {
            radial_values(pl.L2, k * r, REG, rad);
            u_s[0] = u[0]; u_s[1] = u[1]; u_s[2] = u[2];
        }
    }
    __syncthreads();
    const double u[3] = {u_s[0], u_s[1], u_s[2]};
    for (int q = threadIdx.x; q <= pl.L2; q += blockDim.x) {
        if (ok_s) {
            table_column(q, pl.L2, u, rad, pl.refl, tab);
        } else {
            zero_column(q, pl.L2, tab);
            if (REG && q == 0) tab[0] = cd(1.0);
        }
    }
    __syncthreads();
    const int Md = 2 * pl.Hd;
    cd *o = out + (size_t)blk * Md * 2 * pl.Hs;
    for (int e = threadIdx.x; e < pl.Hd * pl.Hs; e += blockDim.x) {
        const int a = e / pl.Hd, b = e - a * pl.Hd;
        const int s0 = pl.starts[e], s1 = pl.starts[e + 1], ns = pl.nsame[e];
        cd same(0.0), cross(0.0);
        for (int t = s0; t < s0 + ns; ++t) same += pl.w[t] * tab[pl.hidx[t]];
        for (int t = s0 + ns; t < s1; ++t) cross += pl.w[t] * tab[pl.hidx[t]];
        o[(size_t)(2 * a) * Md + 2 * b] = cd(same.re - cross.im, same.im + cross.re);          // pol 1
        o[(size_t)(2 * a + 1) * Md + 2 * b + 1] = cd(same.re + cross.im, same.im - cross.re);  // pol 0
    }
}

int launch_translation_rect(const RectPlan &pl, double k, int64_t count, const double *disp, int regular, cd *out,
                            hipStream_t s) {
    const size_t blk = (size_t)4 * pl.Hd * pl.Hs;
    TG_HIP(hipMemsetAsync(out, 0, sizeof(cd) * blk * count, s));
    if (count == 0) return TG_OK;
    const size_t smem = sizeof(cd) * (pl.T + pl.L2 + 1);
    if (regular)
        k_translation_rect<true><<<(unsigned)count, 256, smem, s>>>(pl, k, disp, out);
    else
        k_translation_rect<false><<<(unsigned)count, 256, smem, s>>>(pl, k, disp, out);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// --------------------------------------------------- matrix-free coupling (K2) --

// y_i = sum_{j != i} C_ij x_j (or the regular overlap R with identity self blocks),
// recomputing every pair block in FP64.  Workgroup (destination i, chunk of sources);
// thread (g, b) owns destination row b and source columns a = g (mod G).  Partial
// sums per chunk go to `part` [chunk][col][h][nh] and are added in chunk order.
template <int NC, bool REG>
__global__ __launch_bounds__(256) void k_coupling_mf(DevPlan pl, double k, const double *pos, int N, int cols,
                                                      const cd *x, int64_t n, int64_t nh, int chunk, cd *part) {
    extern __shared__ cd sm[];
    const int T = pl.T, L2 = pl.L2, H = pl.H;
    const int G = blockDim.x / H;
    cd *tab = sm;
    cd *rad = tab + T;
    cd *xs = rad + (L2 + 1);  // [cols][2][H]
    __shared__ double su[3];
    __shared__ int flag;
    const int tid = threadIdx.x;
    const int g = tid / H, b = tid - g * H;
    const int i = blockIdx.x;
    const int j0 = blockIdx.y * chunk, j1 = min(N, j0 + chunk);
    cd acc[NC][2];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) { acc[cc][0] = cd(0.0); acc[cc][1] = cd(0.0); }
    for (int j = j0; j < j1; ++j) {
        __syncthreads();
        if (tid == 0) {
            double d[3], r, u[3];
            for (int a = 0; a < 3; ++a) d[a] = pos[3 * i + a] - pos[3 * j + a];
            int f = (i == j || !unit_direction(d, r, u)) ? 1 : 2;
            if (f == 2) {
                radial_values(L2, k * r, REG, rad);
                su[0] = u[0]; su[1] = u[1]; su[2] = u[2];
            }
            flag = f;
        }
        for (int t = tid; t < cols * 2 * H; t += blockDim.x) {
            const int cc = t / (2 * H), rem = t - cc * 2 * H, h = rem / H, a = rem - h * H;
            xs[t] = x[(int64_t)cc * n + (int64_t)h * nh + (int64_t)j * H + a];
        }
        __syncthreads();
        if (flag == 1) {
            if (REG && g == 0 && tid < G * H) {
#pragma unroll
                for (int cc = 0; cc < NC; ++cc)
                    if (cc < cols) {
                        acc[cc][0] += xs[cc * 2 * H + b];
                        acc[cc][1] += xs[cc * 2 * H + H + b];
                    }
            }
            continue;
        }
        for (int q = tid; q <= L2; q += blockDim.x) {
            double u[3] = {su[0], su[1], su[2]};
            table_column(q, L2, u, rad, pl.refl, tab);
        }
        __syncthreads();
        if (tid < G * H) {
            for (int a = g; a < H; a += G) {
                cd cp, cm;
                entry_values(pl, a * H + b, tab, cp, cm);
#pragma unroll
                for (int cc = 0; cc < NC; ++cc)
                    if (cc < cols) {
                        acc[cc][0] += cp * xs[cc * 2 * H + a];
                        acc[cc][1] += cm * xs[cc * 2 * H + H + a];
                    }
            }
        }
    }
    // reduce the G column groups in LDS (fixed order)
    __syncthreads();
    cd *red = sm;  // [G][cols][2][H]
    if (tid < G * H) {
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
            if (cc < cols) {
                red[((g * cols + cc) * 2 + 0) * H + b] = acc[cc][0];
                red[((g * cols + cc) * 2 + 1) * H + b] = acc[cc][1];
            }
    }
    __syncthreads();
    for (int t = tid; t < cols * 2 * H; t += blockDim.x) {
        cd s(0.0);
        for (int gg = 0; gg < G; ++gg) s += red[gg * cols * 2 * H + t];
        const int cc = t / (2 * H), rem = t - cc * 2 * H, h = rem / H, bb = rem - h * H;
        part[((int64_t)blockIdx.y * cols + cc) * n + (int64_t)h * nh + (int64_t)i * H + bb] = s;
    }
}

// out[t] = sum_c part[c * count + t], c = 0..nchunks-1 in order.
template <class PT>
__global__ void k_sum_chunks(const PT *part, int nchunks, int64_t count, cd *out) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < nchunks; ++c) {
        PT v = part[(int64_t)c * count + t];
        re += (double)v.x;
        im += (double)v.y;
    }
    out[t] = cd(re, im);
}

static int grid1(int64_t count, int block = 256) { return (int)((count + block - 1) / block); }

// ------------------------------------------- matrix-free coupling with pair tables --
//
// The FP64 coupling is the plan contraction C_ij[b, a] = sum_t w_t tab_ij[h_t] (same
// and cross kinds, combined per helicity), so
//   y_i[b] = sum_j sum_a C_ij[b, a] x_j[a] = sum_{a, t} w_t U_{h_t}[a][i],
//   U_h[q][i] = sum_j TAB_h[i][j] X[q][j],
// with TAB_h[i][j] = tab_ij[h] the singular harmonic table of the pair (zero for
// i = j) and q = (column, helicity, source mode a).  U is one complex GEMM per
// harmonic index h (M = N destinations, K = N sources) on MFMA, over only the source
// modes a that some term of h uses (57% of all (h, a) at lmax = 6: the Gaunt
// selection rules |l_a - l_b| <= p <= l_a + l_b and m_b = m_a + q leave the rest
// zero), and the plan contraction that follows reads U once per term.

// TAB[h][j][i] = z_p(k r) P_p^q e^{i q phi} at d = pos_i - pos_j for every ordered pair,
// z_p = h_p (singular coupling) or j_p (REG: regular overlap, whose self blocks are the
// identity: the harmonic table at the origin is 1 at p = q = 0).  One thread per
// (i, j) with i fastest, so each table entry is a coalesced store across the wave.
// The singular table h_p(x) P_p^q e^{i q phi} comes from singular_table_put
// (tg_pairtab.h): the table_column recurrences with the Hankel upward recurrence run
// alongside each order q, the same operations in the same order, so the same values.
template <bool REG>
__global__ void k_pair_tables(DevPlan pl, double k, const double *pos, int N, cd *tab, cf *tab32) {
    const int L2 = pl.L2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * N) return;
    const int j = (int)(t / N), i = (int)(t - (int64_t)j * N);
    const int64_t NN = (int64_t)N * N;
    const int64_t o = (int64_t)j * N + i;
    // the FP64 entry and, when asked for, its c64 rounding (the FP32 MFMA tables)
    auto put = [&](int h, cd v) {
        stream_store(tab + (int64_t)h * NN + o, v);
        if (tab32) stream_store(tab32 + (int64_t)h * NN + o, cf((float)v.re, (float)v.im));
    };
    double d[3], r, u[3];
    for (int a = 0; a < 3; ++a) d[a] = pos[3 * i + a] - pos[3 * j + a];
    if ((!REG && i == j) || !unit_direction(d, r, u)) {
        for (int q = 0; q <= L2; ++q) zero_column_put(q, L2, put);
        if (REG) put(0, cd(1.0));
        return;
    }
    if (!REG) {
        singular_table_put(L2, k * r, u, pl.refl, put);
        return;
    }
    cd rad[2 * kMaxLmax + 2];
    radial_values(L2, k * r, REG, rad);
    for (int q = 0; q <= L2; ++q) table_column_put(q, L2, u, rad, pl.refl, put);
}

// ------------------------------------------------------------ tangent tables --
// Directional derivative of the harmonic table z_p(k r) S_p^q(d / r) along a
// displacement tangent dd and a wavenumber tangent dk (sw/plan.rs:552-595,
// TranslationPlan::tangent_table): with S the solid harmonic at the unit vector u and
// g its Cartesian gradient there,
//   d/dd_a = z_p'(k r) k u_a S + z_p(k r) / r (g_a - p u_a S),   d/dk = z_p'(k r) r S,
// contracted with (dd, dk).  Negative orders: reflection factor times the conjugate
// of S and g (z_p is not conjugated).  One order q at a time: the ladder P_q^q, then
// the degree recurrence of S and of its gradient (r^2 = 1 on the unit sphere, with
// gradient 2 u), as in the pullback's derivative table.
struct SolidGrad {
    cd s, g0, g1, g2;
};

// z(p) returns z_p and z_{p+1} at x = k r for p = q..L2 (called in ascending p).
template <class Radial, class Put>
__device__ __forceinline__ void tangent_column_put(int q, int L2, double k, double r, const double u[3],
                                                   const double dd[3], double dk, const double *refl,
                                                   const SolidGrad &ladder, Radial radial, Put put) {
    const double ud = u[0] * dd[0] + u[1] * dd[1] + u[2] * dd[2];
    const double inv_x = 1.0 / (k * r);
    cd s = ladder.s, g0 = ladder.g0, g1 = ladder.g1, g2 = ladder.g2;
    cd ps(0.0), pg0(0.0), pg1(0.0), pg2(0.0);
    for (int p = q; p <= L2; ++p) {
        if (p > q) {
            const double a = (double)(2 * p - 1) / (double)(p - q);
            const double b = (double)(p + q - 1) / (double)(p - q);
            const cd ns = (a * u[2]) * s - b * ps;
            const cd n0 = a * (u[2] * g0) - b * (pg0 + (2.0 * u[0]) * ps);
            const cd n1 = a * (u[2] * g1) - b * (pg1 + (2.0 * u[1]) * ps);
            const cd n2 = a * (u[2] * g2 + s) - b * (pg2 + (2.0 * u[2]) * ps);
            ps = s; pg0 = g0; pg1 = g1; pg2 = g2;
            s = ns; g0 = n0; g1 = n1; g2 = n2;
        }
        cd zp, zn;
        radial(p, zp, zn);
        const cd zd = ((double)p * inv_x) * zp - zn;  // z_p'(x)
        const cd radial_part = zd * (k * ud + r * dk);
        const cd ang = zp * (1.0 / r);
        const cd gd = dd[0] * g0 + dd[1] * g1 + dd[2] * g2;
        put(p * p + p + q, radial_part * s + ang * (gd - ((double)p * ud) * s));
        if (q > 0) {
            const double f = refl[p * p + p - q];
            const cd sn = f * conj(s), gn = f * conj(gd);
            put(p * p + p - q, radial_part * sn + ang * (gn - ((double)p * ud) * sn));
        }
    }
}

__device__ __forceinline__ void ladder_step(int q, const double u[3], SolidGrad &l) {
    const cd xy(u[0], u[1]);
    const double f = -(double)(2 * q - 1);
    l.g0 = f * (l.g0 * xy + l.s);
    l.g1 = f * (l.g1 * xy + l.s * cd(0.0, 1.0));
    l.g2 = f * (l.g2 * xy);
    l.s = f * (l.s * xy);
}

// TAN[h][j][i] = tangent of TAB[h][j][i] for the position tangents dpos (3 N) and dk:
// dd = dpos_i - dpos_j.  Singular tables (zero for i = j); one thread per (i, j), i
// fastest; the Hankel upward recurrence runs alongside each order (no indexed array).
__global__ void k_pair_tangent_tables(DevPlan pl, double k, const double *pos, const double *dpos, double dk, int N,
                                      cd *tab) {
    const int L2 = pl.L2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * N) return;
    const int j = (int)(t / N), i = (int)(t - (int64_t)j * N);
    const int64_t NN = (int64_t)N * N;
    const int64_t o = (int64_t)j * N + i;
    auto put = [&](int h, cd v) { stream_store(tab + (int64_t)h * NN + o, v); };
    double d[3], r, u[3], dd[3];
    for (int a = 0; a < 3; ++a) {
        d[a] = pos[3 * i + a] - pos[3 * j + a];
        dd[a] = dpos[3 * i + a] - dpos[3 * j + a];
    }
    if (i == j || !unit_direction(d, r, u)) {
        for (int q = 0; q <= L2; ++q) zero_column_put(q, L2, put);
        return;
    }
    const double x = k * r, inv = 1.0 / x, sx = sin(x), cx_ = cos(x);
    SolidGrad ladder{cd(1.0), cd(0.0), cd(0.0), cd(0.0)};
    for (int q = 0; q <= L2; ++q) {
        if (q > 0) ladder_step(q, u, ladder);
        // h_{p} and h_{p+1} from h_{-1}, h_0 upward, advanced to p = q first
        cd hprev(cx_ * inv, sx * inv), hcur(sx * inv, -cx_ * inv);
        for (int p = 0; p < q; ++p) {
            const cd hn = ((double)(2 * p + 1) * inv) * hcur - hprev;
            hprev = hcur;
            hcur = hn;
        }
        int at = q;
        cd hnext = ((double)(2 * q + 1) * inv) * hcur - hprev;
        auto radial = [&](int p, cd &zp, cd &zn) {
            while (at < p) {
                const cd hn = ((double)(2 * at + 3) * inv) * hnext - hcur;
                hcur = hnext;
                hnext = hn;
                ++at;
            }
            zp = hcur;
            zn = hnext;
        };
        tangent_column_put(q, L2, k, r, u, dd, dk, pl.refl, ladder, radial, put);
    }
}

int launch_pair_tangent_tables(tg_cluster *c, const double *dpos_dev, double dk, cd *tab) {
    const int64_t NN = (int64_t)c->N * c->N;
    k_pair_tangent_tables<<<grid1(NN), 256, 0, c->stream>>>(c->plan, c->k0, c->b_pos.as<double>(), dpos_dev, dk, c->N,
                                                            tab);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// Tangent tables of a displacement list (one thread per displacement, [count][T]):
// singular or regular radial functions from an array of z_0 .. z_{L2+1}.  A regular
// table at the origin is constant (identity block), so its tangent is zero there.
template <bool REG>
__global__ void k_tangent_tables_list(DevPlan pl, double k, int64_t count, const double *disp, const double *tdisp,
                                      double dk, cd *tab) {
    const int L2 = pl.L2, T = pl.T;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    cd *out = tab + e * T;
    auto put = [&](int h, cd v) { out[h] = v; };
    double d[3], r, u[3], dd[3];
    for (int a = 0; a < 3; ++a) {
        d[a] = disp[3 * e + a];
        dd[a] = tdisp[3 * e + a];
    }
    if (!unit_direction(d, r, u)) {
        for (int q = 0; q <= L2; ++q) zero_column_put(q, L2, put);
        return;
    }
    cd rad[2 * kMaxLmax + 2];
    radial_values(L2 + 1, k * r, REG, rad);
    SolidGrad ladder{cd(1.0), cd(0.0), cd(0.0), cd(0.0)};
    for (int q = 0; q <= L2; ++q) {
        if (q > 0) ladder_step(q, u, ladder);
        auto radial = [&](int p, cd &zp, cd &zn) {
            zp = rad[p];
            zn = rad[p + 1];
        };
        tangent_column_put(q, L2, k, r, u, dd, dk, pl.refl, ladder, radial, put);
    }
}

// M x M blocks from per-displacement tables [count][T] (TranslationPlan::evaluate_table,
// sw/plan.rs:224-251), one workgroup per block, the table in LDS.
__global__ __launch_bounds__(256) void k_blocks_from_tables(DevPlan pl, const cd *tab, double2 *out) {
    extern __shared__ cd tl[];
    const int T = pl.T, H = pl.H, M = 2 * H;
    const int64_t e = blockIdx.x;
    for (int t = threadIdx.x; t < T; t += blockDim.x) tl[t] = tab[e * T + t];
    __syncthreads();
    OutBlocks ob{out, M};
    for (int en = threadIdx.x; en < H * H; en += blockDim.x) {
        const int a = en / H, b = en - a * H;
        cd vp, vm;
        entry_values(pl, en, tl, vp, vm);
        ob.store(0, (int)e, a, b, vp, vm);
    }
}

int launch_translation_tangent(int lmax, const DevPlan &pl, double k, int64_t count, const double *disp,
                               const double *tdisp, double dk, int regular, cd *tables, cd *outp, hipStream_t s) {
    const int M = 2 * lmax * (lmax + 2);
    TG_TRY(launch_zero_f64((double *)outp, (int64_t)2 * M * M * count, s));
    if (count == 0) return TG_OK;
    if (regular)
        k_tangent_tables_list<true><<<grid1(count, 64), 64, 0, s>>>(pl, k, count, disp, tdisp, dk, tables);
    else
        k_tangent_tables_list<false><<<grid1(count, 64), 64, 0, s>>>(pl, k, count, disp, tdisp, dk, tables);
    TG_HIP(hipGetLastError());
    k_blocks_from_tables<<<(unsigned)count, 256, sizeof(cd) * pl.T, s>>>(pl, tables, (double2 *)outp);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// FP64 pair tables rounded to c64, for the FP32 MFMA coupling.
__global__ void k_tables_to_c64(const cd *tab, int64_t count, cf *out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) out[t] = cf((float)tab[t].re, (float)tab[t].im);
}

// Pair tables are kept when one set takes at most c->table_limit bytes.
bool tables_fit(const tg_cluster *c) {
    return (int64_t)sizeof(cd) * c->N * c->N * c->plan.T <= c->table_limit;
}

// log2 of the bound |h_p(k0 dmin)| sqrt((2p)!), maximised over p, of every pair table
// entry (-inf for a single sphere)
static double log2_table_bound(const tg_cluster *c) {
    if (!(c->dmin < INFINITY)) return -INFINITY;  // a single sphere: no pair tables
    const double x = c->k0 * c->dmin, inv = 1.0 / x;
    std::complex<double> hprev(std::cos(x) * inv, std::sin(x) * inv), hcur(std::sin(x) * inv, -std::cos(x) * inv);
    double log2max = -INFINITY, lfact = 0.0;  // lfact = ln (2p)!
    for (int p = 0; p <= c->plan.L2; ++p) {
        if (p > 0) lfact += std::log((double)(2 * p - 1)) + std::log((double)(2 * p));
        log2max = std::max(log2max, std::log2(std::abs(hcur)) + 0.5 * lfact / std::log(2.0));
        const std::complex<double> hnext = ((double)(2 * p + 1) * inv) * hcur - hprev;
        hprev = hcur;
        hcur = hnext;
    }
    return log2max;
}

bool c64_tables_in_range(const tg_cluster *c) {
    const double b = log2_table_bound(c);
    return b == -INFINITY || (std::isfinite(b) && b + std::log2((double)c->N) < 110.0);
}

bool fp32_tables_in_range(const tg_cluster *c) {
    const double b = log2_table_bound(c);
    return b == -INFINITY || (std::isfinite(b) && b < 110.0);
}

// The FP64 singular (or regular) tables, and with c64 also their c64 rounding: built
// together in one pass when neither exists yet.
static int ensure_pair_tables(tg_cluster *c, bool regular, bool c64) {
    TG_TRY(tables_wait(c));
    bool &have = regular ? c->have_tablesR : c->have_tables;
    DevBuf &buf = regular ? c->b_tabR : c->b_tab;
    const bool need32 = c64 && !regular && !c->have_tables32;
    if (have && !need32) return TG_OK;
    const int64_t NN = (int64_t)c->N * c->N, count = NN * c->plan.T;
    if (need32) TG_TRY(c->b_tab32.ensure(sizeof(cf) * (size_t)count));
    cf *t32 = need32 ? c->b_tab32.as<cf>() : nullptr;
    if (have) {
        k_tables_to_c64<<<grid1(count), 256, 0, c->stream>>>(buf.as<cd>(), count, t32);
    } else {
        TG_TRY(buf.ensure(sizeof(cd) * (size_t)count));
        if (regular)
            k_pair_tables<true><<<grid1(NN), 256, 0, c->stream>>>(c->plan, c->k0, c->b_pos.as<double>(), c->N,
                                                                  buf.as<cd>(), nullptr);
        else
            k_pair_tables<false><<<grid1(NN), 256, 0, c->stream>>>(c->plan, c->k0, c->b_pos.as<double>(), c->N,
                                                                   buf.as<cd>(), t32);
    }
    TG_HIP(hipGetLastError());
    have = true;
    if (need32) c->have_tables32 = true;
    return TG_OK;
}

int tables_wait(tg_cluster *c) {
    c->split_pending = false;
    if (!c->tables_pending) return TG_OK;
    c->tables_pending = false;
    TG_HIP(hipStreamWaitEvent(c->stream, c->side_done, 0));
    return TG_OK;
}

int split_wait(tg_cluster *c) {
    if (!c->split_pending) return TG_OK;
    c->split_pending = false;
    TG_HIP(hipStreamWaitEvent(c->stream, c->side_split, 0));
    return TG_OK;
}

int side_sync(tg_cluster *c) {
    c->tables_pending = false;
    c->split_pending = false;
    if (c->side) TG_HIP(hipStreamSynchronize(c->side));
    return TG_OK;
}

// The inner coupling's tables are built on the side stream, after the work already on
// the cluster's stream, while the cluster's stream sets up the preconditioner (its
// kernels and host round trips): the split FP16 tables first, straight from the
// recurrences (the first inner apply waits for them only), then the FP64 tables of the
// refinement residuals (their first user waits for them).
int prefetch_inner_tables(tg_cluster *c) {
    if (c->tables_pending || c->have_tablesh || c->inner_gemm != TG_INNER_SPLIT16 || !tables_fit(c) ||
        !c64_tables_in_range(c))
        return TG_OK;
    if (!c->side) {
        TG_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        TG_HIP(hipEventCreateWithFlags(&c->side_start, hipEventDisableTiming));
        TG_HIP(hipEventCreateWithFlags(&c->side_split, hipEventDisableTiming));
        TG_HIP(hipEventCreateWithFlags(&c->side_done, hipEventDisableTiming));
    }
    TG_HIP(hipEventRecord(c->side_start, c->stream));
    TG_HIP(hipStreamWaitEvent(c->side, c->side_start, 0));
    hipStream_t main_stream = c->stream;
    c->stream = c->side;
    int st = ensure_split_tables(c);
    if (hipEventRecord(c->side_split, c->side) != hipSuccess && !st) {
        set_error("hipEventRecord failed on the side stream");
        st = TG_DEVICE_ERROR;
    }
    if (!st) st = ensure_pair_tables(c, false, false);
    c->stream = main_stream;
    if (hipEventRecord(c->side_done, c->side) != hipSuccess && !st) {
        set_error("hipEventRecord failed on the side stream");
        st = TG_DEVICE_ERROR;
    }
    c->split_pending = true;
    c->tables_pending = true;  // even after an error: the cluster's stream still waits
    return st;
}

// X[q][j] with q = (column cc, helicity h, mode a): the GEMM's right operand, in the
// GEMM's precision.
template <class T>
__global__ void k_pack_sources(const cd *x, int64_t n, int64_t nh, int N, int H, int cols, cx<T> *X) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t Q = (int64_t)cols * 2 * H;
    if (t >= Q * N) return;
    const int64_t q = t / N;
    const int j = (int)(t - q * N);
    const int cc = (int)(q / (2 * H)), rem = (int)(q - (int64_t)cc * 2 * H), h = rem / H, a = rem - h * H;
    const cd v = x[(int64_t)cc * n + (int64_t)h * nh + (int64_t)j * H + a];
    X[t] = cx<T>((T)v.re, (T)v.im);
}


// U[g][gpos[soff[t] + r]][i] = sum_j TAB_t[i][j] X[g H + smode[soff[t] + r]][j] for
// every harmonic index t, column group g = (column, helicity) < G and subset rank r
// (columns stored in the contraction's group order): one
// BM x BN tile per workgroup (64 x 64, or 128 x 32: a harmonic's subset is 13-48
// modes per column group, and 32-wide column tiles pad less).  Tiles are ordered
// (t, row tile, column tile), column
// tile fastest (tiles[t] column tiles before t), and the 1-D grid is remapped so that
// each XCD (workgroups are dealt to the 8 XCDs round-robin) runs a contiguous range
// of them: the column tiles sharing one BM-row panel of TAB_t then run together on
// one XCD and read the panel through its L2.
template <class T, bool KTAIL, int BM, int BN>
__global__ __launch_bounds__(256, 2) void k_zgemm_subsets(const cx<T> *TAB, const cx<T> *X, cx<T> *U, int N, int H,
                                                          int S, int G, int T_, const int32_t *tiles,
                                                          const int32_t *soff, const int32_t *smode,
                                                          const int32_t *gpos) {
    constexpr int WC = BN / 32;
    const int nb = gridDim.x, bid = blockIdx.x;
    const int xcd = bid & 7, q8 = nb >> 3, r8 = nb & 7;
    const int L = xcd * q8 + min(xcd, r8) + (bid >> 3);
    const int R = (N + BM - 1) / BM;  // row tiles
    int lo = 0, hi = T_ - 1;      // last t with tiles[t] R <= L
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tiles[mid] * R <= L) lo = mid; else hi = mid - 1;
    }
    const int t = lo;
    const int nct = tiles[t + 1] - tiles[t], local = L - tiles[t] * R;
    const int rt = local / nct, ct = local - rt * nct;
    const int s0 = soff[t], nt = soff[t + 1] - s0, ncol = G * nt;
    const int64_t c0 = (int64_t)ct * BN, r0 = (int64_t)rt * BM;
    __shared__ TileLds<T, BM, BN> sm;
    Tile3M<T> acc;
    auto bcol = [=](int64_t c) {
        const int q = (int)min(c, (int64_t)ncol - 1), g = q / nt, r = q - g * nt;
        return X + ((int64_t)g * H + smode[s0 + r]) * N;
    };
    zgemm_tile_3m_cols<T, KTAIL, BM, BN>(TAB + (int64_t)t * N * N, N, bcol, r0, c0, N, N, sm, acc);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave / WC) * 32, wn = (wave % WC) * 32, rr = lane & 15;
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int q = (int)c0 + wn + b * 16 + mfma_out_row<T>(lane, v);
            if (q >= ncol) continue;
            const int g = q / nt, r = q - g * nt;
            cx<T> *Uc = U + ((int64_t)g * S + gpos[s0 + r]) * N;
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int64_t gr = r0 + wm + a * 16 + rr;
                const T t1 = acc.t1[a][b][v], t2 = acc.t2[a][b][v], t3 = acc.t3[a][b][v];
                if (gr < N) Uc[gr] = cx<T>(t1 - t2, t3 - t1 - t2);
            }
        }
}

// The plan contraction y[cc][h][i H + b] = sum over the terms (a, t) of destination
// mode b of w_t U[(cc, h)][column of (t, a)][i], same kind plus or minus i times cross
// kind (h = 0: pol 1, +; h = 1: pol 0, -).  Column s of U (harmonic (p, q), source mode
// (l, m)) feeds only the modes (lambda, mu = m - q), so a workgroup of 4 waves per
// (column group g, 64 spheres, order mu) reads each of its columns once per sphere (U
// is read once in all, against about three times per value when each destination mode
// reads its own terms) and keeps the same- and cross-kind sums of every lambda in
// registers.  U stores the columns in group order (column s at
// position gpos[s]), so an order's columns are contiguous; wave w takes columns w,
// w + 4, ... of them (wave-uniform, so the weights are scalar loads); the four waves'
// sums are added in LDS in wave order.  LB >= lmax bounds the register arrays.
template <class T, int LB, int NW = 4, int UN = 8>
__global__ __launch_bounds__(NW * 64) void k_tab_contract_mu(DevPlan pl, const cx<T> *__restrict__ U, int N, int H,
                                                             int64_t n, int64_t nh, cd *__restrict__ y) {
    const int lmax = pl.lmax, ng = 2 * lmax + 1, NI = (N + 63) / 64;
    const int nb = gridDim.x, bid = blockIdx.x;
    const int xcd = bid & 7, q8 = nb >> 3, r8 = nb & 7;
    const int L = xcd * q8 + min(xcd, r8) + (bid >> 3);
    // sphere blocks fastest: the workgroups reading one column's N values run together
    const int ib = L % NI, rest = L / NI, mg = rest % ng, g = rest / ng;
    // wave index made provably uniform, so that the column list and weights are scalar
    // loads
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = ib * 64 + lane, ic = min(i, N - 1);
    const int k0 = pl.gstart[mg], k1 = pl.gstart[mg + 1];
    const cx<T> *Ug = U + (int64_t)g * pl.S * N + ic;
    double sr[LB], si[LB], cr[LB], ci[LB];
#pragma unroll
    for (int l = 0; l < LB; ++l) sr[l] = si[l] = cr[l] = ci[l] = 0.0;
    // U holds the columns in group order (the GEMMs store column s at gpos[s]), so the
    // order's columns are one contiguous block of k1 - k0 columns
    for (int k = k0 + wave; k < k1; k += NW * UN) {
        cx<T> u[UN];
#pragma unroll
        for (int v = 0; v < UN; ++v) u[v] = Ug[(int64_t)min(k + NW * v, k1 - 1) * N];
#pragma unroll
        for (int v = 0; v < UN; ++v) {
            if (k + NW * v >= k1) break;
            const double *wk = pl.gw + (int64_t)(k + NW * v) * lmax * 2;
            const double ur = (double)u[v].re, ui = (double)u[v].im;
#pragma unroll
            for (int l = 0; l < LB; ++l) {
                if (l >= lmax) break;
                const double ws = wk[2 * l], wc = wk[2 * l + 1];
                sr[l] = fma(ws, ur, sr[l]);
                si[l] = fma(ws, ui, si[l]);
                cr[l] = fma(wc, ur, cr[l]);
                ci[l] = fma(wc, ui, ci[l]);
            }
        }
    }
    // one lambda at a time through LDS (every wave reaches every barrier: lmax and mu are
    // uniform)
    __shared__ double red[NW - 1][4][64];
    const int mu = mg - lmax, cc = g >> 1, h = g & 1;
    cd *yo = y + (int64_t)cc * n + (int64_t)h * nh + (int64_t)ic * H;
#pragma unroll
    for (int l = 0; l < LB; ++l) {
        const int lam = l + 1;
        if (lam > lmax) break;
        if (lam < (mu < 0 ? -mu : mu)) continue;
        if (wave > 0) {
            red[wave - 1][0][lane] = sr[l];
            red[wave - 1][1][lane] = si[l];
            red[wave - 1][2][lane] = cr[l];
            red[wave - 1][3][lane] = ci[l];
        }
        __syncthreads();
        if (wave == 0 && i < N) {
            double a0 = sr[l], a1 = si[l], a2 = cr[l], a3 = ci[l];
#pragma unroll
            for (int w = 0; w < NW - 1; ++w) {
                a0 += red[w][0][lane];
                a1 += red[w][1][lane];
                a2 += red[w][2][lane];
                a3 += red[w][3][lane];
            }
            yo[lam * lam - 1 + mu + lam] = h == 0 ? cd(a0 - a3, a1 + a2) : cd(a0 + a3, a1 - a2);
        }
        __syncthreads();
    }
}

// 4 waves of 4 loads in flight: at N = 1024, lmax 6, 4 columns 112 us (FP64 U) and
// 91 us (c64 U), against 134 / 93 us for 4 x 8, 158 / 108 for 8 x 8, 155 / 124 for
// 4 x 16, 115 / 93 for 4 x 2, 126 / 105 for 8 x 4 and 171 / 164 for 2 x 8
constexpr int kContractWaves = 4, kContractLoads = 4;

// y = the plan contraction of U for nc right-hand sides (2 nc column groups)
template <class T>
static int launch_contract(tg_cluster *c, const cx<T> *U, int nc, cd *y) {
    const int N = c->N, H = c->H, lmax = c->plan.lmax;
    const unsigned grid = (unsigned)(((N + 63) / 64) * (2 * lmax + 1) * 2 * nc);
#define TG_CONTRACT(LB)                                                                                           \
    k_tab_contract_mu<T, LB, kContractWaves, kContractLoads><<<grid, kContractWaves * 64, 0, c->stream>>>(        \
        c->plan, U, N, H, c->n, c->nh, y)
    if (lmax <= 2) TG_CONTRACT(2);
    else if (lmax <= 4) TG_CONTRACT(4);
    else if (lmax <= 6) TG_CONTRACT(6);
    else if (lmax <= 8) TG_CONTRACT(8);
    else TG_CONTRACT(12);
#undef TG_CONTRACT
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// The tables in the GEMM precision: FP64 pair tables, or their c64 rounding (built
// once from the FP64 tables and kept with them).
template <class T> static const cx<T> *pair_tables(tg_cluster *c, bool regular);
template <> const cd *pair_tables<double>(tg_cluster *c, bool regular) {
    return (regular ? c->b_tabR : c->b_tab).as<cd>();
}
template <> const cf *pair_tables<float>(tg_cluster *c, bool) { return c->b_tab32.as<cf>(); }

// C x over a given table set TAB (GEMM precision R): pack the sources, one subset GEMM
// per harmonic index, the FP64 plan contraction.
template <class R>
static int coupling_on_tables(tg_cluster *c, const cx<R> *TAB, const cd *x, cd *y, int cols, const char *name,
                              const char *gname);

template <class R>
static int coupling_tables(tg_cluster *c, const cd *x, cd *y, int cols, bool regular) {
    TG_TRY(ensure_pair_tables(c, regular, sizeof(R) == 4));
    const char *name = regular ? "overlap_mf" : sizeof(R) == 8 ? "coupling_mf" : "coupling_mf32";
    const char *gname = regular ? "gemm_overlap" : sizeof(R) == 8 ? "gemm_f64" : "gemm_c64";
    return coupling_on_tables<R>(c, pair_tables<R>(c, regular), x, y, cols, name, gname);
}

int coupling_tables_f64(tg_cluster *c, const cd *TAB, const cd *x, cd *y, int cols) {
    return coupling_on_tables<double>(c, TAB, x, y, cols, "coupling_tangent", "gemm_tangent");
}

template <class R>
static int coupling_on_tables(tg_cluster *c, const cx<R> *TAB, const cd *x, cd *y, int cols, const char *name,
                              const char *gname) {
    const int N = c->N, H = c->H, T = c->plan.T;  // T: harmonic indices
    const int S = c->plan.S;
    // right-hand sides per pass: the packed sources and U of a pass within 512 MB
    // (at least 8, at most kGemmCols)
    const size_t per_col = sizeof(cx<R>) * ((size_t)2 * H * N + (size_t)2 * S * N);
    const int cmax = (int)std::min<size_t>(kGemmCols, std::max<size_t>(8, ((size_t)512 << 20) / per_col));
    const int64_t Qmax = (int64_t)std::min(cols, cmax) * 2 * H;
    TG_TRY(c->ws[0].ensure(sizeof(cx<R>) * ((size_t)Qmax * N + (size_t)std::min(cols, cmax) * 2 * S * N)));
    cx<R> *X = c->ws[0].as<cx<R>>();
    cx<R> *U = X + (size_t)Qmax * N;
    Timer::Rec rec;
    TG_TRY(c->timer.begin(name, c->stream, rec));
    for (int c0 = 0; c0 < cols; c0 += cmax) {
        const int nc = std::min(cmax, cols - c0);
        const int64_t Q = (int64_t)nc * 2 * H;
        k_pack_sources<R><<<grid1(Q * N), 256, 0, c->stream>>>(x + (int64_t)c0 * c->n, c->n, c->nh, N, H, nc, X);
        // FP32: 64 x 64 tiles (w = 0); FP64: 128 x 32 (w = 1)
        constexpr int w = sizeof(R) == 8 ? 1 : 0, BN = kTileWidths[w], BM = 64 * 64 / BN;
        const int32_t *tiles = c->plan.stiles + (size_t)(w * kGemmCols + nc - 1) * (T + 1);
        const unsigned grid = (unsigned)c->gemm_ntiles[(size_t)w * kGemmCols + nc - 1] * (unsigned)((N + BM - 1) / BM);
        Timer::Rec grec;
        TG_TRY(c->timer.begin(gname, c->stream, grec));
        if (N % 16)
            k_zgemm_subsets<R, true, BM, BN><<<grid, 256, 0, c->stream>>>(TAB, X, U, N, H, S, 2 * nc, T, tiles,
                                                                         c->plan.soff, c->plan.smode, c->plan.gpos);
        else
            k_zgemm_subsets<R, false, BM, BN><<<grid, 256, 0, c->stream>>>(TAB, X, U, N, H, S, 2 * nc, T, tiles,
                                                                          c->plan.soff, c->plan.smode, c->plan.gpos);
        TG_HIP(hipGetLastError());
        TG_TRY(c->timer.end(gname, c->stream, grec));
        TG_TRY(launch_contract<R>(c, U, nc, y + (int64_t)c0 * c->n));
    }
    TG_TRY(c->timer.end(name, c->stream, rec));
    return TG_OK;
}

// C x with the c64 tables on FP32 MFMA (accumulated in FP32, contracted in FP64): the
// inner operator of mixed-precision solves without a stored coupling.  On split-FP16
// tables the hi halves alone are multiplied when hi_only is set or the current inner
// step allows the error (relaxation).
int coupling_matrix_free32(tg_cluster *c, const cd *x, cd *y, int cols, bool hi_only) {
    if (!tables_fit(c) || !c64_tables_in_range(c)) return coupling_matrix_free(c, x, y, cols, 0);
    if (c->inner_gemm == TG_INNER_SPLIT16) {
        hi_only = hi_only || (c->relax_threshold > 0.0 && c->relax_allowed >= c->relax_threshold);
        return coupling_tables_split16(c, x, y, cols, hi_only);
    }
    return coupling_tables<float>(c, x, y, cols, false);
}

int launch_tab_contract_c64(tg_cluster *c, const cf *U, int nc, cd *y) { return launch_contract<float>(c, U, nc, y); }

int coupling_matrix_free(tg_cluster *c, const cd *x, cd *y, int cols, int regular) {
    // pair tables and GEMMs while the tables fit the cluster's limit; beyond it the
    // translation coefficients are recomputed per pair in every apply
    if (tables_fit(c)) return coupling_tables<double>(c, x, y, cols, regular != 0);
    const int H = c->H;
    const int G = 256 / H;
    const int block = G * H;
    // chunks of sources: enough workgroups to fill the chip
    int chunk = c->N;
    while ((int64_t)c->N * ((c->N + chunk - 1) / chunk) < 4096 && chunk > 8) chunk = (chunk + 1) / 2;
    const int nchunks = (c->N + chunk - 1) / chunk;
    const int64_t count = (int64_t)cols * c->n;
    TG_TRY(c->ws[0].ensure(sizeof(cd) * count * nchunks));
    cd *part = c->ws[0].as<cd>();
    size_t smem = sizeof(cd) * (c->plan.T + c->plan.L2 + 1 + (size_t)cols * 2 * H);
    size_t red = sizeof(cd) * (size_t)G * cols * 2 * H;
    if (red > smem) smem = red;
    dim3 grid(c->N, nchunks);
    Timer::Rec rec;
    const char *name = regular ? "overlap_mf" : "coupling_mf";
    TG_TRY(c->timer.begin(name, c->stream, rec));
    for (int c0 = 0; c0 < cols; c0 += 8) {
        int nc = cols - c0 < 8 ? cols - c0 : 8;
        const cd *xx = x + (int64_t)c0 * c->n;
        cd *pp = part + (int64_t)c0 * c->n;  // chunk stride below uses `cols`
        (void)pp;
        // process column block [c0, c0+nc) with its own partial buffer layout
        cd *pb = part;
#define TG_MF(NCV)                                                                                        \
    if (regular)                                                                                          \
        k_coupling_mf<NCV, true><<<grid, block, smem, c->stream>>>(c->plan, c->k0, c->b_pos.as<double>(), \
                                                                   c->N, nc, xx, c->n, c->nh, chunk, pb); \
    else                                                                                                  \
        k_coupling_mf<NCV, false><<<grid, block, smem, c->stream>>>(c->plan, c->k0, c->b_pos.as<double>(), \
                                                                    c->N, nc, xx, c->n, c->nh, chunk, pb);
        if (nc == 1) { TG_MF(1) } else if (nc == 2) { TG_MF(2) } else if (nc <= 4) { TG_MF(4) } else { TG_MF(8) }
#undef TG_MF
        TG_HIP(hipGetLastError());
        int64_t cnt = (int64_t)nc * c->n;
        k_sum_chunks<double2><<<grid1(cnt), 256, 0, c->stream>>>((const double2 *)pb, nchunks, cnt,
                                                                  y + (int64_t)c0 * c->n);
        TG_HIP(hipGetLastError());
    }
    TG_TRY(c->timer.end(name, c->stream, rec));
    return TG_OK;
}

// ------------------------------------------------------ stored matvec (GEMV) --

template <class S> struct acc_of;
template <> struct acc_of<float2> { using T = float; using V = float2; };
template <> struct acc_of<double2> { using T = double; using V = double2; };

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <class S> struct vec16;
template <> struct vec16<float2> { using T = f32x4; static constexpr int R = 2; };   // two c64 rows
template <> struct vec16<double2> { using T = f64x2; static constexpr int R = 1; };  // one c128 row

// x packed for the matvec: xp[h][k][NC] in the accumulation precision, zero beyond
// `cols`, so each workgroup copies its chunk with contiguous 16-byte loads.
template <class AV, int NC>
__global__ void k_pack_x(const cd *x, int64_t n, int64_t nh, int cols, AV *xp) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * nh * NC) return;
    const int cc = (int)(t % NC);
    const int64_t hk = t / NC;  // h * nh + k
    AV v;
    if (cc < cols) {
        const cd xv = x[(int64_t)cc * n + hk];
        v.x = xv.re;
        v.y = xv.im;
    } else {
        v.x = 0;
        v.y = 0;
    }
    xp[t] = v;
}

// y_h = C_h x_h for the stored split coupling, both helicities (grid.z), split over
// column chunks (grid.y).  Each thread owns R consecutive rows and streams the
// chunk's columns with one 16-byte non-temporal load per column, issued in batches
// of 8 so that eight loads are in flight per wave.  The packed x is read with
// wave-uniform addresses (scalar loads), so the kernel needs no LDS and runs at full
// occupancy.  Partial sums [kc][col][h][nh] are added in chunk order.
template <class S, int NC>
__global__ __launch_bounds__(256) void k_matvec_stored(const S *C, int64_t nh,
                                                        const typename acc_of<S>::V *__restrict__ xp, int64_t n,
                                                        int cols, int kchunk, typename acc_of<S>::V *part) {
    using A = typename acc_of<S>::T;
    using AV = typename acc_of<S>::V;
    constexpr int R = vec16<S>::R;
    constexpr int B = 8;
    using V16 = typename vec16<S>::T;
    const int h = blockIdx.z;
    const S *Ch = C + (int64_t)h * nh * nh;
    const int64_t c0 = (int64_t)blockIdx.y * kchunk;
    const int64_t c1 = min(nh, c0 + kchunk);
    const int kc = (int)(c1 - c0);
    const AV *xs = xp + ((int64_t)h * nh + c0) * NC;
    const int64_t r0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * R;
    if (r0 >= nh) return;
    A acc[NC][R][2];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc)
#pragma unroll
        for (int r = 0; r < R; ++r) acc[cc][r][0] = acc[cc][r][1] = 0;
    auto mac = [&](const V16 &v, int k) {
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
            const AV xv = xs[k * NC + cc];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const A re = (A)v[2 * r], im = (A)v[2 * r + 1];
                acc[cc][r][0] += re * xv.x - im * xv.y;
                acc[cc][r][1] += re * xv.y + im * xv.x;
            }
        }
    };
    if (nh % R == 0) {
        const V16 *p = reinterpret_cast<const V16 *>(Ch + c0 * nh + r0);
        const int64_t stride = nh / R;
        int k = 0;
        for (; k + B <= kc; k += B) {
            V16 v[B];
#pragma unroll
            for (int u = 0; u < B; ++u) v[u] = __builtin_nontemporal_load(p + (int64_t)(k + u) * stride);
#pragma unroll
            for (int u = 0; u < B; ++u) mac(v[u], k + u);
        }
        for (; k < kc; ++k) mac(__builtin_nontemporal_load(p + (int64_t)k * stride), k);
    } else {
        const S *p = Ch + c0 * nh + r0;
        for (int k = 0; k < kc; ++k)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (r0 + r >= nh) break;
                const S v0 = p[(int64_t)k * nh + r];
#pragma unroll
                for (int cc = 0; cc < NC; ++cc) {
                    const AV xv = xs[k * NC + cc];
                    acc[cc][r][0] += (A)v0.x * xv.x - (A)v0.y * xv.y;
                    acc[cc][r][1] += (A)v0.x * xv.y + (A)v0.y * xv.x;
                }
            }
    }
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
        if (cc >= cols) break;
        AV *o = part + ((int64_t)blockIdx.y * cols + cc) * n + (int64_t)h * nh + r0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (r0 + r >= nh) break;
            AV v;
            v.x = acc[cc][r][0];
            v.y = acc[cc][r][1];
            o[r] = v;
        }
    }
}

template <class S>
static int matvec_stored_t(tg_cluster *c, const cd *x, cd *y, int cols) {
    using AV = typename acc_of<S>::V;
    const int64_t nh = c->nh;
    // column chunk per workgroup: long enough to amortise the partial sums, short
    // enough to give every CU several workgroups (measured on config-5 shapes)
    const int kchunk = sizeof(AV) == 8 ? 4096 : 2048;
    const int nk = (int)((nh + kchunk - 1) / kchunk);
    TG_TRY(c->ws[1].ensure(sizeof(AV) * (size_t)nk * 8 * c->n + sizeof(AV) * 8 * (size_t)c->n));
    AV *part = c->ws[1].as<AV>();
    AV *xp = part + (size_t)nk * 8 * c->n;  // packed x: [2][nh][NC]
    const S *C = c->b_C.as<S>();
    Timer::Rec rec;
    TG_TRY(c->timer.begin("matvec_stored", c->stream, rec));
    for (int c0 = 0; c0 < cols; c0 += 8) {
        const int nc = cols - c0 < 8 ? cols - c0 : 8;
        const int64_t rows_per_block = 256 * vec16<S>::R;
        dim3 grid((unsigned)((nh + rows_per_block - 1) / rows_per_block), nk, 2);
        const cd *xx = x + (int64_t)c0 * c->n;
#define TG_MV(NCV)                                                                                            \
    k_pack_x<AV, NCV><<<grid1(2 * nh * NCV), 256, 0, c->stream>>>(xx, c->n, nh, nc, xp);                   \
    k_matvec_stored<S, NCV><<<grid, 256, 0, c->stream>>>(C, nh, xp, c->n, nc, kchunk, part);
        if (nc == 1) { TG_MV(1) } else if (nc == 2) { TG_MV(2) } else if (nc <= 4) { TG_MV(4) } else { TG_MV(8) }
#undef TG_MV
        TG_HIP(hipGetLastError());
        const int64_t cnt = (int64_t)nc * c->n;
        k_sum_chunks<AV><<<grid1(cnt), 256, 0, c->stream>>>(part, nk, cnt, y + (int64_t)c0 * c->n);
        TG_HIP(hipGetLastError());
    }
    TG_TRY(c->timer.end("matvec_stored", c->stream, rec));
    return TG_OK;
}

int coupling_stored(tg_cluster *c, const cd *x, cd *y, int cols) {
    if (c->storage == TG_STORE_C64) return matvec_stored_t<float2>(c, x, y, cols);
    if (c->storage == TG_STORE_C128) return matvec_stored_t<double2>(c, x, y, cols);
    set_error("no stored coupling: call tg_cluster_assemble first");
    return TG_INVALID_INPUT;
}

// ------------------------------------------------------------- local T --

// y = T x (or T^H x); with sub != NULL, y = sub - T x.  One thread per
// (column, sphere, (l, m)); the 2x2 block of degree l mixes the two helicities.
__global__ void k_local(const cd *mie, int lmax, int N, int H, int64_t n, int64_t nh, int cols, const cd *x, cd *y,
                        int adjoint, const cd *sub) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)cols * nh) return;
    const int64_t cc = t / nh, k = t - cc * nh;
    const int p = (int)(k / H), a = (int)(k - (int64_t)p * H);
    const int l = (int)sqrt((double)(a + 1) + 0.5);
    const cd *blk = mie + (size_t)(p * lmax + (l - 1)) * 4;
    cd t00 = blk[0], t01 = blk[1], t10 = blk[2], t11 = blk[3];
    if (adjoint) {
        cd s = t01;
        t00 = conj(t00); t11 = conj(t11);
        t01 = conj(t10); t10 = conj(s);
    }
    const int64_t ip = cc * n + k, im = ip + nh;
    cd xp = x[ip], xm = x[im];
    cd yp = t00 * xp + t01 * xm, ym = t10 * xp + t11 * xm;
    if (sub) {
        yp = sub[ip] - yp;
        ym = sub[im] - ym;
    }
    y[ip] = yp;
    y[im] = ym;
}

int local_apply_blocks(tg_cluster *c, const cd *blocks, const cd *x, cd *y, int cols, int adjoint, const cd *sub) {
    int64_t count = (int64_t)cols * c->nh;
    k_local<<<grid1(count), 256, 0, c->stream>>>(blocks, c->lmax, c->N, c->H, c->n, c->nh, cols, x, y, adjoint, sub);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

int local_apply(tg_cluster *c, const cd *x, cd *y, int cols, int adjoint, const cd *sub) {
    if (c->particles) return local_apply_particles(c, x, y, cols, adjoint, sub);
    int64_t count = (int64_t)cols * c->nh;
    k_local<<<grid1(count), 256, 0, c->stream>>>(c->b_mie.as<cd>(), c->lmax, c->N, c->H, c->n, c->nh, cols, x, y,
                                                 adjoint, sub);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// --------------------------------------------------------- permutations --

__global__ void k_ref_split(const cd *in, cd *out, int M, int H, int64_t n, int64_t nh, int64_t count, int to_split) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int64_t cc = t / n, r = t - cc * n;  // r: reference index
    const int64_t p = r / M;
    const int mode = (int)(r - p * M);
    const int a = mode >> 1, h = mode & 1;
    const int64_t s = cc * n + (int64_t)h * nh + p * H + a;
    if (to_split)
        out[s] = in[t];
    else
        out[t] = in[s];
}

int ref_to_split(tg_cluster *c, const cd *ref, cd *split, int cols) {
    int64_t count = (int64_t)cols * c->n;
    k_ref_split<<<grid1(count), 256, 0, c->stream>>>(ref, split, c->M, c->H, c->n, c->nh, count, 1);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

int split_to_ref(tg_cluster *c, const cd *split, cd *ref, int cols) {
    int64_t count = (int64_t)cols * c->n;
    k_ref_split<<<grid1(count), 256, 0, c->stream>>>(split, ref, c->M, c->H, c->n, c->nh, count, 0);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// y = P conj(x) with (P x)[p, (l, m)] = (-1)^m x[p, (l, -m)], per helicity.  The
// coupling satisfies C_h = P C_h^T P (reciprocity), so C_h^H = P conj(C_h) P.
__global__ void k_reflect_conj(const cd *x, cd *y, int H, int64_t count) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int64_t base = t - (t % H);
    const int a = (int)(t % H);
    const int l = (int)sqrt((double)(a + 1) + 0.5);
    const int m = a + 1 - l * l - l;
    const int ar = l * l + l - m - 1;
    cd v = conj(x[base + ar]);
    y[t] = (m & 1) ? -v : v;
}

int reflect_conj(tg_cluster *c, const cd *x, cd *y, int cols) {
    int64_t count = (int64_t)cols * c->n;
    k_reflect_conj<<<grid1(count), 256, 0, c->stream>>>(x, y, c->H, count);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// ------------------------------------------------------------ plane wave --

// Regular local coefficients of a unit plane wave (helicity pol) in the split
// layout: 2 sqrt(pi (2l+1)/(l(l+1))) sqrt((l-m)!/(l+m)!) i^l e^{-i m phi}
// (tau + (2 pol - 1) pi) e^{i k.r_p}, with the angular functions from
// F_l^m = P_l^m / sin^|m| so the poles need no division.
__global__ void k_plane_wave(int N, int H, int64_t nh, const double *pos, double kx, double ky, double kz, int pol,
                             cd *out) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * nh) return;
    const int h = (int)(t / nh);
    const int64_t k = t - (int64_t)h * nh;
    const int p = (int)(k / H), a = (int)(k - (int64_t)p * H);
    const int wave_h = pol == 1 ? 0 : 1;
    if (h != wave_h) {
        out[t] = cd(0.0);
        return;
    }
    const int l = (int)sqrt((double)(a + 1) + 0.5);
    const int m = a + 1 - l * l - l;
    const double kk = sqrt(kx * kx + ky * ky + kz * kz);
    const double tr = sqrt(kx * kx + ky * ky);
    const bool axis = tr == 0.0;
    const double cx_ = axis ? 0.0 : kx / tr, sy_ = axis ? 0.0 : ky / tr;
    const double cth = axis ? (kz / kk >= 0.0 ? 1.0 : -1.0) : kz / kk;
    const double sth = tr / kk;
    // F_l^mm and F_{l-1}^mm for mm = (m == 0 ? 1 : m)
    const int mm = m == 0 ? 1 : m;
    const int am = mm < 0 ? -mm : mm;
    double f = 1.0;
    if (mm >= 0)
        for (int i = 1; i <= mm; ++i) f *= -(double)(2 * i - 1);
    else
        for (int i = 1; i <= -mm; ++i) f *= 1.0 / (double)(2 * i);
    double fl = 0.0, flm1 = 0.0;
    if (l >= am) {
        double prev = 0.0, cur = f;
        for (int d = am + 1; d <= l; ++d) {
            double next = ((double)(2 * d - 1) * cth * cur - (double)(d + mm - 1) * prev) / (double)(d - mm);
            prev = cur;
            cur = next;
        }
        fl = cur;
        flm1 = (l > am) ? prev : 0.0;
    }
    double pi_v, tau_v;
    if (m == 0) {
        pi_v = 0.0;
        tau_v = sth * fl;
    } else {
        double power = 1.0;
        for (int i = 1; i < am; ++i) power *= sth;
        pi_v = (double)m * power * fl;
        tau_v = power * ((double)l * cth * fl - (double)(l + m) * flm1);
    }
    const double ang = tau_v + (2.0 * pol - 1.0) * pi_v;
    double ratio = 1.0;  // (l-m)!/(l+m)!
    if (m > 0)
        for (int i = l - m + 1; i <= l + m; ++i) ratio /= (double)i;
    else
        for (int i = l + m + 1; i <= l - m; ++i) ratio *= (double)i;
    const double norm = 2.0 * sqrt(M_PI * (2 * l + 1) / (double)(l * (l + 1))) * sqrt(ratio);
    cd az(1.0);
    if (!axis) {
        cd e(cx_, -sy_);
        int am2 = m < 0 ? -m : m;
        for (int i = 0; i < am2; ++i) az = az * e;
        if (m < 0) az = conj(az);
    }
    const double phase = kx * pos[3 * p] + ky * pos[3 * p + 1] + kz * pos[3 * p + 2];
    cd v = (norm * ang) * (ipow_(l) * az * cd(cos(phase), sin(phase)));
    out[t] = v;
}

int plane_wave(tg_cluster *c, const double *dir, int pol, cd *out) {
    double nrm = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    double kx = c->k0 * dir[0] / nrm, ky = c->k0 * dir[1] / nrm, kz = c->k0 * dir[2] / nrm;
    k_plane_wave<<<grid1(2 * c->nh), 256, 0, c->stream>>>(c->N, c->H, c->nh, c->b_pos.as<double>(), kx, ky, kz, pol,
                                                          out);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// ------------------------------------------------------------ vector ops --

// conj(V[col][i]) . w[col] for nvec vectors per column: block (x, col) sums its slice
// of k (up to 8 vectors per pass over the slice, w read once per pass; reduced across
// the wave by shuffles and across the 4 waves in LDS) into
// partial[(col nvec + i) gridDim.x + x].
__global__ __launch_bounds__(256) void k_dots(const cd *__restrict__ V, int64_t vstride, int nvec,
                                               const cd *__restrict__ w, int64_t n, int64_t csv, int64_t csw,
                                               cd *__restrict__ partial) {
    constexpr int CH = 8;
    __shared__ double sred[4][CH][2];
    const int col = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const cd *wc = w + col * csw;
    const cd *Vc = V + col * csv;
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t k0 = blockIdx.x * per, k1 = min(n, k0 + per);
    for (int i0 = 0; i0 < nvec; i0 += CH) {
        double re[CH], im[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) re[c] = im[c] = 0.0;
        for (int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
            const cd bv = wc[k];
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                if (i0 + c < nvec) {
                    const cd a = Vc[(int64_t)(i0 + c) * vstride + k];
                    re[c] += a.re * bv.re + a.im * bv.im;
                    im[c] += a.re * bv.im - a.im * bv.re;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < CH; ++c)
            for (int off = 32; off > 0; off >>= 1) {
                re[c] += __shfl_xor(re[c], off);
                im[c] += __shfl_xor(im[c], off);
            }
        if (lane == 0)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                sred[wave][c][0] = re[c];
                sred[wave][c][1] = im[c];
            }
        __syncthreads();
        if (threadIdx.x < CH && i0 + (int)threadIdx.x < nvec) {
            const int c = threadIdx.x;
            const double r = (sred[0][c][0] + sred[1][c][0]) + (sred[2][c][0] + sred[3][c][0]);
            const double m = (sred[0][c][1] + sred[1][c][1]) + (sred[2][c][1] + sred[3][c][1]);
            partial[((int64_t)col * nvec + i0 + c) * gridDim.x + blockIdx.x] = cd(r, m);
        }
        __syncthreads();
    }
}

// out[r] = sum of partial row r (width values), zero where mask[r / nvec] is 0, to
// out_dev and/or host_out[copy_len + r], after copy_src[0 .. copy_len) is copied to
// host_out (workgroup 0).  Rows of at most 16 values: one thread per row, values in
// order; wider rows: one wave per row, lane l adding values l, l + 64, ..., then a
// fixed shuffle tree.
__global__ __launch_bounds__(256) void k_dots_finish(const cd *__restrict__ partial, int width, int rows, int nvec,
                                                      const double *mask, cd *out_dev, cd *host_out,
                                                      const cd *copy_src, int64_t copy_len) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (blockIdx.x == 0)
        for (int64_t k = threadIdx.x; k < copy_len; k += blockDim.x) host_out[k] = copy_src[k];
    if (width <= 16) {
        const int r = blockIdx.x * blockDim.x + threadIdx.x;
        if (r >= rows) return;
        const cd *pr = partial + (int64_t)r * width;
        cd s(0.0);
        for (int b = 0; b < width; ++b) s += pr[b];
        if (mask && mask[r / nvec] == 0.0) s = cd(0.0);
        if (out_dev) out_dev[r] = s;
        if (host_out) host_out[copy_len + r] = s;
        return;
    }
    for (int r = blockIdx.x * 4 + wave; r < rows; r += 4 * gridDim.x) {
        const cd *pr = partial + (int64_t)r * width;
        double sre = 0.0, sim = 0.0;
        for (int b = lane; b < width; b += 64) {
            const cd v = pr[b];
            sre += v.re;
            sim += v.im;
        }
        for (int off = 32; off > 0; off >>= 1) {
            sre += __shfl_xor(sre, off);
            sim += __shfl_xor(sim, off);
        }
        if (lane == 0) {
            cd s(sre, sim);
            if (mask && mask[r / nvec] == 0.0) s = cd(0.0);
            if (out_dev) out_dev[r] = s;
            if (host_out) host_out[copy_len + r] = s;
        }
    }
}

static int dot_blocks(int64_t n) {
    int64_t b = (n + 2047) / 2048;
    if (b > 256) b = 256;
    if (b < 1) b = 1;
    return (int)b;
}

static int dots_launch(tg_cluster *c, const cd *V, int64_t vstride, int nvec, const cd *w, int64_t n, int cols,
                       int64_t csv, int64_t csw, const double *mask, cd *out_dev, cd *host_out, const cd *copy_src,
                       int64_t copy_len) {
    const int nb = dot_blocks(n);
    // room for 64 vectors per column at once: the Arnoldi basis grows by one vector per
    // iteration, and the scratch should not be reallocated every time
    const size_t cap = (size_t)cols * ((nvec + 63) / 64 * 64);
    TG_TRY(c->ws[4].ensure(sizeof(cd) * cap * nb));
    cd *partial = c->ws[4].as<cd>();
    k_dots<<<dim3(nb, cols), 256, 0, c->stream>>>(V, vstride, nvec, w, n, csv, csw, partial);
    TG_HIP(hipGetLastError());
    const int rows = cols * nvec;
    const unsigned fblocks = nb <= 16 ? (unsigned)((rows + 255) / 256) : (unsigned)std::min(64, (rows + 3) / 4);
    k_dots_finish<<<fblocks, 256, 0, c->stream>>>(partial, nb, rows, nvec, mask, out_dev, host_out, copy_src,
                                                  copy_len);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

int vec_dots(tg_cluster *c, const cd *V, int64_t vstride, int nvec, const cd *w, int64_t n, int cols, int64_t csv,
             int64_t csw, cd *out_host) {
    const int rows = cols * nvec;
    TG_TRY(c->pin.ensure(sizeof(cd) * (size_t)rows));
    cd *res = c->pin.as<cd>();
    TG_TRY(dots_launch(c, V, vstride, nvec, w, n, cols, csv, csw, nullptr, nullptr, res, nullptr, 0));
    TG_HIP(hipStreamSynchronize(c->stream));
    std::copy(res, res + rows, out_host);
    return TG_OK;
}

int vec_dots_dev(tg_cluster *c, const cd *V, int64_t vstride, int nvec, const cd *w, int64_t n, int cols,
                 int64_t csv, int64_t csw, cd *out_dev, const double *mask_dev, cd *host_out, const cd *copy_src,
                 int64_t copy_len) {
    return dots_launch(c, V, vstride, nvec, w, n, cols, csv, csw, mask_dev, out_dev, host_out, copy_src, copy_len);
}

int vec_norms(tg_cluster *c, const cd *w, int64_t n, int cols, int64_t colstride, double *out_host) {
    std::vector<cd> d(cols);
    // |w|^2 as the dot of w with itself (nvec = 1, V = w)
    TG_TRY(vec_dots(c, w, 0, 1, w, n, cols, colstride, colstride, d.data()));
    // a NaN stays NaN (fmax would turn it into 0 and pass a non-finite vector as zero)
    for (int i = 0; i < cols; ++i) out_host[i] = d[i].re > 0.0 ? sqrt(d[i].re) : d[i].re == d[i].re ? 0.0 : d[i].re;
    return TG_OK;
}

// w[col] += sign * sum_i coef[col * nvec + i] V[col][i]
__global__ void k_update(const cd *V, int64_t vstride, int nvec, const cd *coef, cd *w, int64_t n, int64_t csv,
                         int64_t csw, double sign) {
    const int col = blockIdx.y;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    cd acc(0.0);
    for (int i = 0; i < nvec; ++i) acc += coef[col * nvec + i] * V[col * csv + (int64_t)i * vstride + k];
    w[col * csw + k] += sign * acc;
}

int vec_update(hipStream_t s, const cd *V, int64_t vstride, int nvec, const cd *coef_dev, cd *w, int64_t n, int cols,
               int64_t csv, int64_t csw, double sign) {
    k_update<<<dim3(grid1(n), cols), 256, 0, s>>>(V, vstride, nvec, coef_dev, w, n, csv, csw, sign);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

__global__ void k_axpby(cd alpha, const cd *x, cd beta, cd *y, int64_t count) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    y[t] = alpha * x[t] + beta * y[t];
}

int vec_axpby(hipStream_t s, cd alpha, const cd *x, cd beta, cd *y, int64_t count) {
    k_axpby<<<grid1(count), 256, 0, s>>>(alpha, x, beta, y, count);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

__global__ void k_scale_cols(const double *scale, const cd *x, cd *y, int64_t n) {
    const int col = blockIdx.y;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    y[col * n + k] = scale[col] * x[col * n + k];
}

int vec_scale_cols(hipStream_t s, const double *scale_dev, const cd *x, cd *y, int64_t n, int cols) {
    k_scale_cols<<<dim3(grid1(n), cols), 256, 0, s>>>(scale_dev, x, y, n);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// The same with the scales and masks passed by value (no host-to-device copy):
// y[col] = v.s[col] x[col], and v (the scales, then the masks) written to `store` for
// the kernels after it.
__global__ void k_scale_cols_v(ColScales v, double *store, const cd *x, cd *y, int64_t n, int cols) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 2 * cols)
        store[threadIdx.x] = threadIdx.x < cols ? v.s[threadIdx.x] : v.m[threadIdx.x - cols];
    const int col = blockIdx.y;
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    y[col * n + k] = v.s[col] * x[col * n + k];
}

int vec_scale_cols_v(hipStream_t s, const ColScales &v, double *store, const cd *x, cd *y, int64_t n, int cols) {
    if (cols > ColScales::kMax) {
        set_error("vec_scale_cols_v: at most %d columns", ColScales::kMax);
        return TG_INVALID_INPUT;
    }
    k_scale_cols_v<<<dim3(grid1(n), cols), 256, 0, s>>>(v, store, x, y, n, cols);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

}  // namespace tg
