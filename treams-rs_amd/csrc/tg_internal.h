// Internal declarations of the gfx950 cluster-coupling library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/treams_gpu.h"
#include "tg_complex.h"
#include "tg_plan.h"

namespace tg {

void set_error(const char *fmt, ...);

#define TG_HIP(expr)                                                                              \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess) {                                                                   \
            ::tg::set_error("HIP error %s at %s:%d (%s)", hipGetErrorString(_e), __FILE__, __LINE__, \
                            #expr);                                                               \
            return _e == hipErrorOutOfMemory ? TG_OUT_OF_MEMORY : TG_DEVICE_ERROR;                 \
        }                                                                                         \
    } while (0)

#define TG_TRY(expr)                   \
    do {                               \
        int _s = (expr);               \
        if (_s != TG_OK) return _s;    \
    } while (0)

constexpr int kGemmCols = 64;  // most right-hand sides per matrix-free GEMM pass
// smallest allowed relative operator error at which an inner split-FP16 apply drops the
// lo halves (tg_cluster_set_inner_relaxation): about three times the hi-only coupling's
// measured relative error (2.8e-4 to 3.0e-4)
constexpr double kRelaxThresholdDefault = 1e-3;
// Column-tile widths of the matrix-free GEMMs: 64 x 64 output tiles in FP32 (the
// fewest loads per MFMA), 128 x 32 in FP64 (less column padding).
constexpr int kTileWidths[2] = {64, 32};

// Device view of the translation plan.
struct DevPlan {
    int H, L2, T;                // modes per helicity, largest harmonic degree, table length
    const int32_t *starts;       // H*H + 1
    const uint8_t *nsame;        // H*H
    const double *w;             // terms
    const uint16_t *hidx;        // terms
    const double *refl;          // T
    // transposed plan for pullbacks: for each harmonic index, its (entry, weight, kind)
    const int32_t *tstarts;      // T + 1
    const int32_t *tentry;       // entry index e (kind in bit 31)
    const double *tw;            // weight
    // source-mode subsets of the matrix-free GEMMs: harmonic index t multiplies only the
    // S_t source modes a that some term (b, a, t) uses, listed in smode[soff[t] ..
    // soff[t + 1]); term k reads subset column ucol[k] (= soff[hidx[k]] + rank of a)
    int S;                       // soff[T]
    const int32_t *soff;         // T + 1
    const int32_t *smode;        // S
    const int32_t *ucol;         // terms
    // GEMM column tiles of width kTileWidths[w] before harmonic t for G = 2 cols column
    // groups: stiles[((w kGemmCols) + cols - 1) (T + 1) + t], cols = 1 .. kGemmCols
    const int32_t *stiles;
    // the subset columns by destination order: column s (harmonic (p, q), source mode
    // (l, m)) feeds only destination modes (lambda, mu = m - q).  The GEMMs store column
    // s at position k = gpos[s]; group mu + lmax holds positions gstart[.] ..
    // gstart[. + 1) (columns ascending), and position k carries the weights
    // gw[(k lmax + lambda - 1) 2 + kind] (kind 0 same, 1 cross; 0 where no term)
    int lmax;
    const int32_t *gstart, *gpos;              // 2 lmax + 2, S
    const double *gw;                          // S lmax 2
};

// Growable device buffer.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    template <class T> T *as() const { return reinterpret_cast<T *>(ptr); }
};

// While one of these is alive on a thread, DevBuf::release marks the blocks it returns
// to the cache clean (no work queued on them): the releasing code has synchronised
// every stream that used them.
struct CleanReleases {
    bool prev;
    CleanReleases();
    ~CleanReleases();
};

// Pinned host memory (hipHostMalloc): kernels write it, the host reads it after the
// stream is synchronised, with no copy in between.
struct HostBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    template <class T> T *as() const { return reinterpret_cast<T *>(ptr); }
};

struct Timer {
    bool enabled = false;
    struct Rec { hipEvent_t a, b; };
    std::map<std::string, std::vector<Rec>> pending;
    std::map<std::string, std::pair<double, int64_t>> totals;
    std::vector<hipEvent_t> spare;  // events of collected records, reused by begin
    int begin(const char *name, hipStream_t s, Rec &r);
    int end(const char *name, hipStream_t s, Rec &r);
    int collect();
    void destroy_events();  // after collect: the spare events
};

// Device view of a rectangular translation plan (destination degrees up to one lmax,
// source degrees up to another); entry e = a Hd + b, source a outer.
struct RectPlan {
    int Hd, Hs, L2, T;
    const int32_t *starts;
    const uint8_t *nsame;
    const double *w;
    const uint16_t *hidx;
    const double *refl;
};

// Upload HostPlan `hp` (square or rectangular) as a RectPlan view into `buf`.
int upload_rect_plan(const HostPlan &hp, DevBuf &buf, RectPlan &rp, hipStream_t s);

struct Preconditioner;  // defined in tg_precond.hip

}  // namespace tg

struct tg_cluster {
    int lmax = 0, N = 0, M = 0, H = 0;
    int64_t n = 0, nh = 0;
    double k0 = 0;
    hipStream_t stream = nullptr;
    std::vector<double> radii, positions;
    std::vector<cd> eps;
    tg::HostPlan hplan;
    tg::DevPlan plan{};
    std::vector<int> gemm_ntiles;      // column tiles of a GEMM pass: [w kGemmCols + cols - 1] (plan.stiles totals)
    tg::DevBuf b_plan, b_pos, b_mie;   // b_mie: value, d/dsize, d/deps blocks
    int storage = TG_STORE_NONE;
    tg::DevBuf b_C;                    // stored split coupling: [2][nh][nh] (c64 or c128)
    tg::DevBuf b_tab;                  // FP64 singular harmonic tables TAB[h][j][i] of all ordered pairs
    bool have_tables = false;
    tg::DevBuf b_tab32;                // the same tables rounded to c64 (FP32 MFMA coupling)
    bool have_tables32 = false;
    tg::DevBuf b_tabh, b_texp;         // split FP16 tables and their row exponents (FP16 MFMA coupling)
    bool have_tablesh = false;
    int inner_gemm = TG_INNER_SPLIT16; // c64-precision GEMMs of the inner solves (tg_inner_gemm)
    double dmin = INFINITY;            // the closest pair of centres (c64_tables_in_range)
    // relaxed inner applies: a split-FP16 apply multiplies the hi halves only when the
    // error it may make (relax_allowed, relative, set per Arnoldi step by an inner
    // GMRES) is at least relax_threshold; 0 switches relaxation off
    double relax_threshold = tg::kRelaxThresholdDefault;
    double relax_allowed = 0.0;
    // particle clusters: arbitrary per-particle T-matrices in place of the Mie blocks,
    // b_tloc[p][h'][a][h][b] = T_p[(h, b)][(h', a)] (per (p, h') a 2H x H column-major
    // panel; h = 0 the reference's polarisation 1)
    bool particles = false;
    tg::DevBuf b_tloc;
    tg::DevBuf b_tabR;                 // FP64 regular harmonic tables (overlap for cross sections)
    bool have_tablesR = false;
    int64_t table_limit = (int64_t)64 << 30;  // largest pair-table set kept (bytes)
    tg::DevBuf ws[8];                  // scratch
    tg::DevBuf outer;                  // residual, correction and A x of mixed-precision refinement
    tg::HostBuf pin;                   // pinned host results of the Krylov reductions
    // side stream that builds the inner coupling's tables while the preconditioner is
    // set up; table users wait for side_done while tables_pending
    hipStream_t side = nullptr;
    hipEvent_t side_start = nullptr, side_split = nullptr, side_done = nullptr;
    bool tables_pending = false;
    bool split_pending = false;        // the split FP16 tables (side_split) are not waited for yet
    tg::Timer timer;
    tg::Preconditioner *pc = nullptr;
    int pc_kind = TG_PC_NONE;
    struct LearnedState *learned = nullptr;  // near pairs, features, weights (tg_learned.hip)
    struct tg_mlp *mlp = nullptr;            // network attached for TG_PC_LEARNED (not owned)
    int group_size = 4;                      // block-Jacobi group size
};

namespace tg {

// Kernel launchers (tg_kernels.hip).  Vectors are in the helicity-split layout
// [col][h][N*H] (h = 0: pol 1, h = 1: pol 0) unless named "ref".
int launch_mie(tg_cluster *c);
int launch_assemble_split(tg_cluster *c, int storage, void *out);
int launch_assemble_ref(tg_cluster *c, int regular, cd *out);
// A = I - T C in the split order, column-major n x n, c64 or c128 (precision: tg_storage)
int launch_assemble_operator(tg_cluster *c, int precision, void *out);
// the same for a particle cluster (arbitrary T blocks): C assembled in the split order,
// then A = I - T C block row by block row on MFMA (tg_particles.hip)
int assemble_operator_particles(tg_cluster *c, int precision, void *out);
int local_apply_particles(tg_cluster *c, const cd *x, cd *y, int cols, int adjoint, const cd *subtract_from);
// dense block-diagonal T' x with T' as split panels (the layout of b_tloc)
int local_apply_dense(tg_cluster *c, const cd *panels, const cd *x, cd *y, int cols, int adjoint,
                      const cd *subtract_from);
// N reference-layout M x M blocks (device) -> split panels
int tloc_from_ref(tg_cluster *c, const cd *ref, cd *panels);
// per-particle sum_c Y R^H (split-layout Y, R) as reference-layout M x M blocks
int local_cotangent(tg_cluster *c, const cd *Y, const cd *R, int cols, cd *out_ref);
// error for operations that need the spheres' Mie blocks (radii, permittivities)
int require_spheres(const tg_cluster *c, const char *what);
// fields of N expansions (outgoing or regular waves) about the centres pos (device,
// xyz), n = N M rows per column; synchronises the stream (tg_fields.hip)
int launch_field(int N, int lmax, double k, const double *pos, const double *refl, const cd *x, int64_t n, int cols,
                 int64_t npoints, const double *points, bool regular, cd *out, hipStream_t s);
// singular blocks as their two same-helicity H x H parts, each entry a pair of bfloat16
// (real part in the low half), [count][h][a][b]
// (fp32_ok: the tables of these pairs fit FP32; fp32_tables_in_range of their cluster)
int launch_translation_blocks_helicity_bf16(int lmax, const DevPlan &plan, double k, int64_t count,
                                           const double *disp, uint32_t *out, hipStream_t s, bool fp32_ok);
int launch_translation_blocks(int lmax, const DevPlan &plan, double k, int64_t count, const double *disp,
                              int regular, cd *out, hipStream_t s);
// tangent tables of every ordered pair (FP64, the layout of the singular pair tables)
// along position tangents dpos_dev (3 N) and a wavenumber tangent dk
int launch_pair_tangent_tables(tg_cluster *c, const double *dpos_dev, double dk, cd *tab);
// tangent blocks of a displacement list (tables: scratch of count x T values)
int launch_translation_tangent(int lmax, const DevPlan &pl, double k, int64_t count, const double *disp,
                               const double *tdisp, double dk, int regular, cd *tables, cd *out, hipStream_t s);
// p[0 .. count) = 0 by a kernel (no memset node)
int launch_zero_f64(double *p, int64_t count, hipStream_t s);
// count blocks of (2 Hd) x (2 Hs), column-major, destination inner
int launch_translation_rect(const RectPlan &pl, double k, int64_t count, const double *disp, int regular, cd *out,
                            hipStream_t s);
int coupling_matrix_free(tg_cluster *c, const cd *x, cd *y, int cols, int regular);
// whether one FP64 pair-table set fits the cluster's table limit
bool tables_fit(const tg_cluster *c);
// whether the c64-precision inner couplings stay within FP32 range: the largest pair
// table entry is bounded by |h_p(k0 dmin)| sqrt((2p)!) over the degrees p (the bound
// of the split tables' row shifts at the closest pair), and that bound times N must
// leave the FP32 intermediates (c64 tables, the GEMM outputs) below 2^110; beyond it
// the inner applies run in FP64
bool c64_tables_in_range(const tg_cluster *c);
// whether that bound stays below 2^110: FP32 copies of the tables (the LDS tables of the
// c64 assemblies), and the sums of at most 2 lmax + 1 of them times coefficients below 2^5
// that form an assembled entry, cannot overflow
bool fp32_tables_in_range(const tg_cluster *c);
// start the FP64 and split tables of the c64-precision inner coupling on the side
// stream (no-op when built, not fitting or not split); tables_wait makes the
// cluster's stream wait for them, side_sync waits on the host (before freeing them)
int prefetch_inner_tables(tg_cluster *c);
int tables_wait(tg_cluster *c);
int side_sync(tg_cluster *c);
// y = C' x for a caller's FP64 table set TAB[h][j][i] (pair-table GEMMs and the plan
// contraction of the coupling, e.g. on tangent tables)
int coupling_tables_f64(tg_cluster *c, const cd *TAB, const cd *x, cd *y, int cols);
// y = T' x (or T'^H x) with 2x2 blocks [sphere][degree][4] in place of the Mie blocks
int local_apply_blocks(tg_cluster *c, const cd *blocks, const cd *x, cd *y, int cols, int adjoint, const cd *sub);
int coupling_matrix_free32(tg_cluster *c, const cd *x, cd *y, int cols, bool hi_only = false);
int coupling_stored(tg_cluster *c, const cd *x, cd *y, int cols);
// split FP16 coupling (tg_split16.hip) from the FP64 singular tables tab64
int ensure_split_tables(tg_cluster *c);
int coupling_tables_split16(tg_cluster *c, const cd *x, cd *y, int cols, bool hi_only = false);
// the cluster's stream waits for the split tables the side stream builds (tables_wait
// waits for everything it builds)
int split_wait(tg_cluster *c);
// the FP64 plan contraction of a compact c64 U for nc columns into y (tg_kernels.hip)
int launch_tab_contract_c64(tg_cluster *c, const cf *U, int nc, cd *y);
int local_apply(tg_cluster *c, const cd *x, cd *y, int cols, int adjoint, const cd *subtract_from);
int ref_to_split(tg_cluster *c, const cd *ref, cd *split, int cols);
int split_to_ref(tg_cluster *c, const cd *split, cd *ref, int cols);
int reflect_conj(tg_cluster *c, const cd *x, cd *y, int cols);
int plane_wave(tg_cluster *c, const double *dir, int pol, cd *out_split);
int pair_gradients(tg_cluster *c, const cd *u, const cd *v, int cols, int regular, double *grad_pos_dev,
                   double *grad_k_dev);

// Operator y = (I - T C) x or its adjoint in the split layout.
int apply_operator(tg_cluster *c, const cd *x, cd *y, int cols, int adjoint, int op);

// Vector helpers (tg_kernels.hip)
// out[col nvec + i] = sum_k conj(V[col][i][k]) w[col][k] (block partial sums, then one
// workgroup adds them in a fixed order), on the cluster's stream with its scratch
int vec_dots(tg_cluster *c, const cd *V, int64_t vstride, int nvec, const cd *w, int64_t n, int cols,
             int64_t colstride_v, int64_t colstride_w, cd *out_host);
int vec_norms(tg_cluster *c, const cd *w, int64_t n, int cols, int64_t colstride, double *out_host);
// vec_dots into device memory out_dev (may be null) and/or pinned host memory host_out
// (after copy_len values copied from copy_src there first), zeroed for columns whose
// mask_dev entry is 0 (mask_dev may be null); no host synchronisation
int vec_dots_dev(tg_cluster *c, const cd *V, int64_t vstride, int nvec, const cd *w, int64_t n, int cols,
                 int64_t csv, int64_t csw, cd *out_dev, const double *mask_dev, cd *host_out = nullptr,
                 const cd *copy_src = nullptr, int64_t copy_len = 0);
int vec_update(hipStream_t s, const cd *V, int64_t vstride, int nvec, const cd *coef_dev, cd *w, int64_t n, int cols,
               int64_t colstride_v, int64_t colstride_w, double sign);
int vec_axpby(hipStream_t s, cd alpha, const cd *x, cd beta, cd *y, int64_t count);
int vec_scale_cols(hipStream_t s, const double *scale_dev, const cd *x, cd *y, int64_t n, int cols);
// per-column scales and masks of a few columns, passed to a kernel by value
struct ColScales {
    static constexpr int kMax = 8;
    double s[kMax], m[kMax];
};
// y = s x per column with s by value; s and m also written to store[0 .. 2 cols)
int vec_scale_cols_v(hipStream_t s, const ColScales &v, double *store, const cd *x, cd *y, int64_t n, int cols);

}  // namespace tg
