// Batched dense LU on the device (tg_lu.hip): a set of column-major complex matrices
// factored together, one matrix per grid.z, with partial pivoting inside 64-row
// diagonal blocks and MFMA trailing updates; solves on packed FP64 vectors.
#pragma once
#include <vector>

#include "tg_internal.h"

namespace tg {

struct LuSet {
    int precision = TG_STORE_C128;  // TG_STORE_C64 (cx<float>) or TG_STORE_C128 (cx<double>)
    void *A = nullptr;              // base of the matrices (owned by the caller)
    int count = 0, nmax = 0;
    int64_t ntot = 0;               // packed vector length: sum of the dimensions
    std::vector<int> n_h;
    std::vector<int64_t> voff_h;
    DevBuf meta;                    // off[count], n[count], voff[count], piv[ntot], perm[ntot], status[count]
    DevBuf work;                    // 8 ntot FP64 complex values
    DevBuf many;                    // the permuted right-hand sides of many-column solves
    const int64_t *off = nullptr;
    const int *n = nullptr;
    const int64_t *voff = nullptr;
    int *piv = nullptr, *perm = nullptr, *status = nullptr;
    // the look-ahead panel stream and its two events, taken at the first factorisation
    // (from the device's pool, or created) and returned to the pool by luset_release
    hipStream_t panel = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    int panel_dev = 0;
};

// Matrix z: dimension n[z], elements from A + off[z] (column-major, leading dim n[z]).
int luset_init(LuSet &ls, int precision, void *A, const std::vector<int64_t> &off, const std::vector<int> &n);
void luset_release(LuSet &ls);
// Factor in place; TG_SINGULAR on a zero pivot.
int luset_factor(LuSet &ls, hipStream_t s);
// The same for a batch of one with partial pivoting over whole columns (GETRF order).
int luset_factor_gepp(LuSet &ls, hipStream_t s);
// x: `cols` packed columns of length ntot (matrix z's rows at voff[z]); overwritten by
// A^{-1} x (adjoint = 0) or A^{-H} x (adjoint = 1), blockwise.
int luset_solve(LuSet &ls, hipStream_t s, cd *x, int cols, int adjoint);

}  // namespace tg
