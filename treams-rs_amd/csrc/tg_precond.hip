// Block-Jacobi preconditioner over spatial groups of spheres (right preconditioning).
//
// Group operator A_G = I - T_G C_GG (sphere-major, reference mode order inside a
// sphere), factored on the device by the batched LU (all groups in one set of
// launches, MFMA trailing updates); applying it gathers each group's rows into a
// packed vector, runs the batched triangular solves and scatters the result back.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <numeric>
#include <vector>

#include "tg_lu.h"
#include "tg_precond.h"

namespace tg {

int launch_translation_blocks(int lmax, const DevPlan &pl, double k, int64_t count, const double *disp, int regular,
                              cd *outp, hipStream_t s);

struct Preconditioner {
    int kind = TG_PC_NONE;
    int group_size = 4;
    int ngroups = 0, max_m = 0;
    DevBuf mats;      // group operators, then their LU factors, column-major, c128
    DevBuf meta;      // offsets (int64), sizes (int), member offsets (int), members (int)
    LuSet lu;         // the batched factorisation of `mats`
    DevBuf packed;    // gathered group vectors [col][sum of group dimensions]
    const int64_t *d_off = nullptr;
    const int *d_size = nullptr, *d_moff = nullptr, *d_members = nullptr;
    // learned near-field correction (tg_learned.hip)
    void *learned = nullptr;
    // precision the current solve asked for: TG_STORE_C64 (inner solves of mixed
    // precision) lets the learned correction stream c64 copies of its near blocks
    int storage = TG_STORE_C128;
};

std::vector<std::vector<int>> spatial_groups(const std::vector<double> &pos, int N, int size) {
    std::vector<std::vector<int>> out;
    std::vector<int> all(N);
    std::iota(all.begin(), all.end(), 0);
    std::vector<std::vector<int>> stack{all};
    while (!stack.empty()) {
        std::vector<int> g = std::move(stack.back());
        stack.pop_back();
        if ((int)g.size() <= size) {
            out.push_back(g);
            continue;
        }
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int p : g)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], pos[3 * p + a]);
                hi[a] = std::max(hi[a], pos[3 * p + a]);
            }
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (hi[a] - lo[a] > hi[axis] - lo[axis]) axis = a;
        std::stable_sort(g.begin(), g.end(), [&](int x, int y) { return pos[3 * x + axis] < pos[3 * y + axis]; });
        // split so that the left part holds a multiple of `size` spheres where possible
        size_t half = g.size() / 2;
        size_t left = ((half + size - 1) / size) * size;
        if (left >= g.size()) left = half;
        std::vector<int> a(g.begin(), g.begin() + left), b(g.begin() + left, g.end());
        stack.push_back(b);
        stack.push_back(a);
    }
    return out;
}

// A_G[R, C] = delta - sum_k T_i[r, k] C_ij[k, c] for R = (li, r), C = (lj, c), i != j;
// the identity on sphere-diagonal blocks.  pair blocks: [pair][M x M] column-major.
__global__ void k_group_operator(const cd *pairs, const int64_t *pair_off, const int *gsize, const int64_t *mat_off,
                                 const int *moff, const int *members, const cd *mie, int lmax, int M, cd *mats) {
    const int g = blockIdx.y;
    const int s = gsize[g];
    const int m = s * M;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m * m) return;
    const int C = (int)(t / m), R = (int)(t - (int64_t)C * m);
    const int li = R / M, r = R - li * M, lj = C / M, cc = C - lj * M;
    cd v(0.0);
    if (li == lj) {
        v = cd(r == cc ? 1.0 : 0.0);
    } else {
        // pair index within the group: ordered pairs (li, lj), li != lj, row-major over li
        const int pidx = li * (s - 1) + (lj < li ? lj : lj - 1);
        const cd *blk = pairs + (pair_off[g] + pidx) * (int64_t)M * M;
        const int b = r >> 1;
        const int l = (int)sqrt((double)(b + 1) + 0.5);
        const int p = members[moff[g] + li];
        const cd *t2 = mie + (size_t)(p * lmax + (l - 1)) * 4;
        const int rr = r & 1;
        cd tc = t2[rr * 2 + 0] * blk[(int64_t)cc * M + 2 * b] + t2[rr * 2 + 1] * blk[(int64_t)cc * M + 2 * b + 1];
        v = -tc;
    }
    mats[mat_off[g] + t] = v;
}

// Packed group vectors: row (slot M + mode) of a packed column holds mode `mode`
// (reference order) of sphere members[slot]; groups are consecutive slot ranges, so
// group g's rows start at moff[g] M, as the batched LU expects.
__device__ __forceinline__ int64_t packed_to_split(const int *members, int M, int H, int64_t nh, int64_t k) {
    const int slot = (int)(k / M), mode = (int)(k - (int64_t)slot * M);
    return (int64_t)(mode & 1) * nh + (int64_t)members[slot] * H + (mode >> 1);
}

__global__ void k_pc_gather(const int *members, int M, int H, int64_t n, int64_t nh, int64_t ld, int cols,
                            const cd *x, cd *packed) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ld * cols) return;
    const int64_t col = t / ld, k = t - col * ld;
    packed[t] = x[col * n + packed_to_split(members, M, H, nh, k)];
}

__global__ void k_pc_scatter(const int *members, int M, int H, int64_t n, int64_t nh, int64_t ld, int cols,
                             const cd *packed, cd *y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ld * cols) return;
    const int64_t col = t / ld, k = t - col * ld;
    y[col * n + packed_to_split(members, M, H, nh, k)] = packed[t];
}

int learned_setup(tg_cluster *c, Preconditioner *pc);
int learned_apply(tg_cluster *c, Preconditioner *pc, const cd *x, cd *y, int cols, int adjoint, int storage);
void learned_destroy(Preconditioner *pc);

static int build_block_jacobi(tg_cluster *c, Preconditioner *pc) {
    const int M = c->M;
    auto groups = spatial_groups(c->positions, c->N, pc->group_size);
    pc->ngroups = (int)groups.size();
    std::vector<int64_t> mat_off(pc->ngroups + 1, 0), pair_off(pc->ngroups + 1, 0);
    std::vector<int> gsize(pc->ngroups), moff(pc->ngroups + 1, 0), members;
    std::vector<double> disp;
    pc->max_m = 0;
    for (int g = 0; g < pc->ngroups; ++g) {
        const auto &G = groups[g];
        const int s = (int)G.size();
        gsize[g] = s;
        moff[g + 1] = moff[g] + s;
        for (int p : G) members.push_back(p);
        mat_off[g + 1] = mat_off[g] + (int64_t)(s * M) * (s * M);
        pair_off[g + 1] = pair_off[g] + (int64_t)s * (s - 1);
        pc->max_m = std::max(pc->max_m, s * M);
        for (int li = 0; li < s; ++li)
            for (int lj = 0; lj < s; ++lj) {
                if (li == lj) continue;
                for (int a = 0; a < 3; ++a) disp.push_back(c->positions[3 * G[li] + a] - c->positions[3 * G[lj] + a]);
            }
    }
    const int64_t npairs = pair_off[pc->ngroups];
    // meta buffer
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t o_off = 0, o_poff = al(sizeof(int64_t) * (pc->ngroups + 1));
    size_t o_size = al(o_poff + sizeof(int64_t) * (pc->ngroups + 1));
    size_t o_moff = al(o_size + sizeof(int) * pc->ngroups);
    size_t o_mem = al(o_moff + sizeof(int) * (pc->ngroups + 1));
    size_t o_disp = al(o_mem + sizeof(int) * members.size());
    size_t total = al(o_disp + sizeof(double) * std::max<size_t>(disp.size(), 1));
    TG_TRY(pc->meta.ensure(total));
    std::vector<char> host(total, 0);
    memcpy(host.data() + o_off, mat_off.data(), sizeof(int64_t) * (pc->ngroups + 1));
    memcpy(host.data() + o_poff, pair_off.data(), sizeof(int64_t) * (pc->ngroups + 1));
    memcpy(host.data() + o_size, gsize.data(), sizeof(int) * pc->ngroups);
    memcpy(host.data() + o_moff, moff.data(), sizeof(int) * (pc->ngroups + 1));
    memcpy(host.data() + o_mem, members.data(), sizeof(int) * members.size());
    if (!disp.empty()) memcpy(host.data() + o_disp, disp.data(), sizeof(double) * disp.size());
    char *base = pc->meta.as<char>();
    TG_HIP(hipMemcpy(base, host.data(), total, hipMemcpyHostToDevice));
    pc->d_off = (const int64_t *)(base + o_off);
    pc->d_size = (const int *)(base + o_size);
    pc->d_moff = (const int *)(base + o_moff);
    pc->d_members = (const int *)(base + o_mem);
    TG_TRY(pc->mats.ensure(sizeof(cd) * (size_t)mat_off[pc->ngroups]));
    DevBuf pairs;
    TG_TRY(pairs.ensure(sizeof(cd) * (size_t)std::max<int64_t>(npairs, 1) * M * M));
    int st = launch_translation_blocks(c->lmax, c->plan, c->k0, npairs, (const double *)(base + o_disp), 0,
                                       pairs.as<cd>(), c->stream);
    if (!st) {
        int64_t maxel = (int64_t)pc->max_m * pc->max_m;
        dim3 grid((unsigned)((maxel + 255) / 256), pc->ngroups);
        k_group_operator<<<grid, 256, 0, c->stream>>>(pairs.as<cd>(), (const int64_t *)(base + o_poff), pc->d_size,
                                                      pc->d_off, pc->d_moff, pc->d_members, c->b_mie.as<cd>(),
                                                      c->lmax, M, pc->mats.as<cd>());
        if (hipGetLastError() != hipSuccess) st = TG_DEVICE_ERROR;
    }
    if (!st) {
        // every group operator, LU-factored in one batch (rows of group g at moff[g] M)
        std::vector<int> dims(pc->ngroups);
        for (int g = 0; g < pc->ngroups; ++g) dims[g] = gsize[g] * M;
        std::vector<int64_t> offs(mat_off.begin(), mat_off.end() - 1);
        st = luset_init(pc->lu, TG_STORE_C128, pc->mats.ptr, offs, dims);
        if (!st) {
            Timer::Rec rec;
            c->timer.begin("pc_factor", c->stream, rec);
            st = luset_factor(pc->lu, c->stream);
            c->timer.end("pc_factor", c->stream, rec);
            if (st == TG_SINGULAR) set_error("singular group block in the block-Jacobi preconditioner");
        }
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess && !st) st = TG_DEVICE_ERROR;
    pairs.release();
    return st;
}

int ensure_preconditioner(tg_cluster *c, int kind, int storage) {
    TG_TRY(require_spheres(c, "the block-Jacobi and learned preconditioners"));
    if (c->pc && c->pc_kind == kind) {
        c->pc->storage = storage;
        return TG_OK;
    }
    destroy_preconditioner(c);
    if (kind != TG_PC_BLOCK_JACOBI && kind != TG_PC_LEARNED) {
        set_error("unknown preconditioner %d", kind);
        return TG_INVALID_INPUT;
    }
    Preconditioner *pc = new Preconditioner();
    pc->kind = kind;
    pc->group_size = c->group_size;
    pc->storage = storage;
    int st = kind == TG_PC_BLOCK_JACOBI ? build_block_jacobi(c, pc) : learned_setup(c, pc);
    if (st) {
        pc->mats.release();
        pc->meta.release();
        luset_release(pc->lu);
        pc->packed.release();
        learned_destroy(pc);
        delete pc;
        return st;
    }
    c->pc = pc;
    c->pc_kind = kind;
    return TG_OK;
}

int apply_preconditioner(tg_cluster *c, const cd *x, cd *y, int cols, int adjoint) {
    Preconditioner *pc = c->pc;
    if (!pc) {
        set_error("preconditioner not built");
        return TG_INVALID_INPUT;
    }
    Timer::Rec rec;
    TG_TRY(c->timer.begin("pc_apply", c->stream, rec));
    if (pc->kind == TG_PC_LEARNED) {
        TG_TRY(learned_apply(c, pc, x, y, cols, adjoint, pc->storage));
    } else {
        const int64_t ld = pc->lu.ntot;  // = N M: every sphere sits in exactly one group
        TG_TRY(pc->packed.ensure(sizeof(cd) * (size_t)ld * cols));
        cd *pk = pc->packed.as<cd>();
        const unsigned grid = (unsigned)((ld * cols + 255) / 256);
        k_pc_gather<<<grid, 256, 0, c->stream>>>(pc->d_members, c->M, c->H, c->n, c->nh, ld, cols, x, pk);
        TG_HIP(hipGetLastError());
        TG_TRY(luset_solve(pc->lu, c->stream, pk, cols, adjoint));
        k_pc_scatter<<<grid, 256, 0, c->stream>>>(pc->d_members, c->M, c->H, c->n, c->nh, ld, cols, pk, y);
        TG_HIP(hipGetLastError());
    }
    TG_TRY(c->timer.end("pc_apply", c->stream, rec));
    return TG_OK;
}

void destroy_preconditioner(tg_cluster *c) {
    if (!c->pc) return;
    c->pc->mats.release();
    c->pc->meta.release();
    luset_release(c->pc->lu);
    c->pc->packed.release();
    learned_destroy(c->pc);
    delete c->pc;
    c->pc = nullptr;
    c->pc_kind = TG_PC_NONE;
}

}  // namespace tg
