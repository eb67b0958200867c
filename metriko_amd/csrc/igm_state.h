// State kept from an iterative rounding for the sized solves (mk_igm_sized_uv / _grad).
#ifndef METRIKO_AMD_IGM_STATE_H
#define METRIKO_AMD_IGM_STATE_H
#include "mk_internal.h"
#include "metriko/core/common/sparse.h"

// The last LM round's system of an iterative rounding, kept for the sized solves: with the
// per-face target scale sigma (the log-sizing hook: F2 rows of face f times sigma_f),
//   (A0 + wK^2 U^T U) x = length wI^2 X2F^T (sigma .* F2) + wC^2 xPrev + wK^2 U^T r,
// solved by Woodbury against A0 with the round's columns Z = A0^{-1} U^T; sigma = 1 gives that
// round's x* (the rounding's result).  Corner uv = P UFull x with the pipeline's corner map P.
struct mk_igm_state {
    mk_ctx* ctx = nullptr;
    mk_solve_opts o{};
    int k = 0, nfull = 0, l1 = 0, nF = 0, N = 0, r = 0;
    double length = 0.0, wI = 0.0, wK = 0.0, wC = 0.0;
    mk_op* A0 = nullptr;   // owned, with its multigrid hierarchy
    mk_op* X2F = nullptr;  // owned, values times length
    metriko::SprsD UF, X2Ft, P;
    std::vector<double> F2, fixedV, UZ;
    std::vector<int> Urows;
    std::vector<int> pos;  // device index of unknown i (A0, X2F, Z, xprev); empty: the same
    mk::DevBuf<int64_t> uptr;
    mk::DevBuf<int> ucol;
    mk::DevBuf<double> uval, Z, xprev;
    bool has_P = false;
    ~mk_igm_state() {
        if (A0) mk_op_destroy(A0);
        if (X2F) mk_op_destroy(X2F);
    }
};

#endif
