// Particle clusters: arbitrary local T-matrices coupled through the singular
// expansion, the dense interaction I - T C of cluster/particles.rs:65-99 (the sphere
// cluster is the special case of Mie blocks).  The particles share one lmax; the
// caller embeds smaller bases and mode subsets with zero rows and columns (their
// modes then solve to zero, exactly the reduced system).  The coupling, the
// matrix-free applies, GMRES, the dense LU, the refinement and the global T-matrix
// are those of the sphere cluster; only the local operator T differs.
#include "tg_internal.h"
#include "tg_mfma.h"

namespace tg {

// b_tloc from T-matrices in the reference mode order (column-major M x M per
// particle; index 2 a + h, h = 0 for polarisation 1).
__global__ void k_tloc_from_ref(const cd *ref, int64_t count, int M, int H, cd *tloc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int64_t MM = (int64_t)M * M;
    const int64_t p = t / MM;
    const int rem = (int)(t - p * MM), s = rem / M, r = rem - s * M;  // column s, row r
    const int b = r >> 1, h = r & 1, a = s >> 1, hp = s & 1;
    tloc[((p * 2 + hp) * H + a) * 2 * H + h * H + b] = ref[t];
}

// y = T x (T^H x with adjoint), or y = sub - T x; split layout; one thread per
// output entry (cc, h, p, b).  x and y must not overlap.
__global__ void k_local_dense(const cd *__restrict__ tloc, int H, int64_t n, int64_t nh, int cols,
                              const cd *__restrict__ x, cd *__restrict__ y, int adjoint,
                              const cd *__restrict__ sub) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)cols * n) return;
    const int64_t cc = t / n, r = t - cc * n;
    const int h = (int)(r / nh);
    const int64_t k = r - (int64_t)h * nh;
    const int64_t p = k / H;
    const int b = (int)(k - p * H);
    const cd *xp = x + cc * n + p * H;
    cd acc(0.0);
    for (int hp = 0; hp < 2; ++hp)
        for (int a = 0; a < H; ++a) {
            const cd tv = adjoint ? conj(tloc[((p * 2 + h) * H + b) * 2 * H + hp * H + a])
                                  : tloc[((p * 2 + hp) * H + a) * 2 * H + h * H + b];
            acc += tv * xp[(int64_t)hp * nh + a];
        }
    y[t] = sub ? sub[t] - acc : acc;
}

int local_apply_dense(tg_cluster *c, const cd *panels, const cd *x, cd *y, int cols, int adjoint, const cd *sub) {
    const int64_t count = (int64_t)cols * c->n;
    k_local_dense<<<(unsigned)((count + 255) / 256), 256, 0, c->stream>>>(panels, c->H, c->n, c->nh, cols, x, y,
                                                                          adjoint, sub);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

int local_apply_particles(tg_cluster *c, const cd *x, cd *y, int cols, int adjoint, const cd *sub) {
    return local_apply_dense(c, c->b_tloc.as<cd>(), x, y, cols, adjoint, sub);
}

int tloc_from_ref(tg_cluster *c, const cd *ref, cd *panels) {
    const int64_t count = (int64_t)c->N * c->M * c->M;
    k_tloc_from_ref<<<(unsigned)((count + 255) / 256), 256, 0, c->stream>>>(ref, count, c->M, c->H, panels);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

// T-bar_p = sum_c Y_p[:, c] R_p[:, c]^H per particle, in the reference mode order
// (column-major M x M per particle), from split-layout Y and R; one thread per entry.
__global__ void k_local_cotangent(const cd *__restrict__ Y, const cd *__restrict__ R, int N, int M, int H,
                                  int64_t n, int64_t nh, int cols, cd *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t MM = (int64_t)M * M;
    if (t >= (int64_t)N * MM) return;
    const int64_t p = t / MM;
    const int rem = (int)(t - p * MM), s = rem / M, r = rem - s * M;
    const int64_t ir = (int64_t)(r & 1) * nh + p * H + (r >> 1), is = (int64_t)(s & 1) * nh + p * H + (s >> 1);
    cd acc(0.0);
    for (int cc = 0; cc < cols; ++cc) acc += Y[(int64_t)cc * n + ir] * conj(R[(int64_t)cc * n + is]);
    out[t] = acc;
}

int local_cotangent(tg_cluster *c, const cd *Y, const cd *R, int cols, cd *out_ref) {
    const int64_t count = (int64_t)c->N * c->M * c->M;
    k_local_cotangent<<<(unsigned)((count + 255) / 256), 256, 0, c->stream>>>(Y, R, c->N, c->M, c->H, c->n, c->nh,
                                                                              cols, out_ref);
    TG_HIP(hipGetLastError());
    return TG_OK;
}

template <class O> __device__ __forceinline__ O out_value(cd v);
template <> __device__ __forceinline__ cd out_value<cd>(cd v) { return v; }
template <> __device__ __forceinline__ cf out_value<cf>(cd v) { return cf((float)v.re, (float)v.im); }

// A rows of particle p (both helicities), columns of helicity h' (blockIdx.z =
// 2 p + h'): A = I - P Q with P the (2H x H) panel T_p[:, h'] and Q the H rows of
// particle p of C_h' (column-major nh x nh); a 64 x 64 tile of P Q per workgroup in the
// 3M form on MFMA, FP64, stored rounded to O.
template <class O, bool KTAIL>
__global__ __launch_bounds__(256, 2) void k_particle_operator(const cd *tloc, const cd *cpm, int H, int64_t nh,
                                                              int64_t n, O *out) {
    const int64_t p = blockIdx.z >> 1;
    const int hp = blockIdx.z & 1;
    const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
    const cd *P = tloc + (p * 2 + hp) * H * 2 * H;
    const cd *Q = cpm + (int64_t)hp * nh * nh + p * H;
    __shared__ TileLds<double> sm;
    Tile3M<double> acc;
    zgemm_tile_3m<double, KTAIL>(P, 2 * H, Q, nh, r0, c0, 2 * H, nh, H, sm, acc);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32, rr = lane & 15;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t gr = r0 + wm + a * 16 + rr, gc = c0 + wn + b * 16 + mfma_out_row<double>(lane, v);
                if (gr >= 2 * H || gc >= nh) continue;
                const double t1 = acc.t1[a][b][v], t2 = acc.t2[a][b][v], t3 = acc.t3[a][b][v];
                const int h = (int)(gr / H);
                const int64_t row = h * nh + p * H + (gr - (int64_t)h * H), col = hp * nh + gc;
                const cd v0 = cd(row == col ? 1.0 - (t1 - t2) : -(t1 - t2), -(t3 - t1 - t2));
                out[col * n + row] = out_value<O>(v0);
            }
}

int assemble_operator_particles(tg_cluster *c, int precision, void *outp) {
    const int64_t nh = c->nh, n = c->n;
    const int H = c->H;
    DevBuf cpm;  // C in the split order: [2][nh][nh] c128, zero self blocks
    TG_TRY(cpm.ensure(sizeof(cd) * 2 * (size_t)nh * nh));
    TG_HIP(hipMemsetAsync(cpm.ptr, 0, sizeof(cd) * 2 * (size_t)nh * nh, c->stream));
    TG_TRY(launch_assemble_split(c, TG_STORE_C128, cpm.ptr));
    const dim3 grid((unsigned)((nh + 63) / 64), (unsigned)((2 * H + 63) / 64), (unsigned)(2 * c->N));
    const bool tail = H % 16 != 0;
#define TG_PARTOP(O, KT) \
    k_particle_operator<O, KT><<<grid, 256, 0, c->stream>>>(c->b_tloc.as<cd>(), cpm.as<cd>(), H, nh, n, (O *)outp)
    if (precision == TG_STORE_C64) {
        if (tail) TG_PARTOP(cf, true); else TG_PARTOP(cf, false);
    } else {
        if (tail) TG_PARTOP(cd, true); else TG_PARTOP(cd, false);
    }
#undef TG_PARTOP
    TG_HIP(hipGetLastError());
    TG_HIP(hipStreamSynchronize(c->stream));
    cpm.release();
    return TG_OK;
}

int require_spheres(const tg_cluster *c, const char *what) {
    if (c && c->particles) {
        set_error("%s needs sphere T-matrices (radii and permittivities); a particle cluster has arbitrary ones", what);
        return TG_INVALID_INPUT;
    }
    return TG_OK;
}

}  // namespace tg

using namespace tg;

extern "C" {

int tg_cluster_create_particles(int lmax, double k0, int n_particles, const double *positions,
                                const tg_c128 *tmatrices, void *stream, tg_cluster **out) {
    if (!out || !positions || !tmatrices || n_particles < 1) {
        set_error("null argument or no particles");
        return TG_INVALID_INPUT;
    }
    *out = nullptr;
    // cluster/particles.rs:101-123: distinct positions; the placeholder spheres below
    // (no contrast, radius below half the distance to the nearest particle) only carry
    // the geometry through the sphere cluster's set-up
    std::vector<double> radii(n_particles, 1.0);
    for (int i = 0; i < n_particles; ++i) {
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(positions[3 * i + a])) {
                set_error("particle positions must be finite");
                return TG_INVALID_INPUT;
            }
        double dmin = INFINITY;
        for (int j = 0; j < n_particles; ++j) {
            if (j == i) continue;
            double d = 0.0;
            for (int a = 0; a < 3; ++a) d = std::hypot(d, positions[3 * i + a] - positions[3 * j + a]);
            dmin = std::min(dmin, d);
        }
        if (!(dmin > 0.0)) {
            set_error("particle modes must be grouped at distinct positions");
            return TG_INVALID_INPUT;
        }
        if (std::isfinite(dmin)) radii[i] = 0.49 * dmin;
    }
    std::vector<tg_c128> eps(n_particles, tg_c128{1.0, 0.0});
    tg_cluster *c = nullptr;
    TG_TRY(tg_cluster_create(lmax, k0, n_particles, radii.data(), eps.data(), positions, stream, &c));
    c->particles = true;
    const int64_t count = (int64_t)n_particles * c->M * c->M;
    int st = c->b_tloc.ensure(sizeof(cd) * (size_t)count);
    if (!st) {
        k_tloc_from_ref<<<(unsigned)((count + 255) / 256), 256, 0, c->stream>>>((const cd *)tmatrices, count, c->M,
                                                                                c->H, c->b_tloc.as<cd>());
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
            set_error("particle T-matrix upload failed");
            st = TG_DEVICE_ERROR;
        }
    }
    if (st) {
        tg_cluster_destroy(c);
        return st;
    }
    *out = c;
    return TG_OK;
}

}  // extern "C"
