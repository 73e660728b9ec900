// suprema_pdlp.hpp -- header-only C++ glue over the C ABI (suprema_pdlp.h).
//
// The functions are templates over the caller's LP and result types, so the
// same code serves the reference's own types (include/ranged_lp.hpp:10-25,
// include/pdlp.hpp:18-24 of the reference) and this repository's mirror of
// them (ranged_lp.hpp / pdlp.hpp here).  An LP type needs: m(), n(),
// A.row_idx / A.col_idx / A.val (COO), c, rL, rU, l, u (std::vector<double>),
// obj_offset.  A result type needs: x, y, iterations, converged, eps_P,
// eps_D, eps_G, restarts.
//
// Errors: a nonzero C status becomes std::runtime_error carrying
// supr_last_error(), the exception the reference's callers catch
// (src/solve.cpp:100-106, 116-122).
#pragma once
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "suprema_pdlp.h"

namespace suprema {

inline void check(int rc, const char* what) {
    if (rc != SUPR_OK)
        throw std::runtime_error(std::string("HIP PDLP error in ") + what + " -- " + supr_last_error());
}

template <class LP>
supr_lp view(const LP& r) {
    supr_lp v;
    v.m = r.m();
    v.n = r.n();
    v.nnz = (int32_t)r.A.val.size();
    v.row_idx = r.A.row_idx.data();
    v.col_idx = r.A.col_idx.data();
    v.val = r.A.val.data();
    v.c = r.c.data();
    v.rL = r.rL.data();
    v.rU = r.rU.data();
    v.l = r.l.data();
    v.u = r.u.data();
    v.obj_offset = r.obj_offset;
    return v;
}

// Engine mode for the drop-in: SUPREMA_PDLP_MODE = plain | fast64 | mixed
// (default fast64: FP64 arithmetic throughout, as the reference's
// solve_pdhg_gpu; mixed runs FP32 iterates with FP64 residuals, restarts and
// polish and must be asked for).
inline int mode_from_env() {
    const char* m = std::getenv("SUPREMA_PDLP_MODE");
    if (!m) return SUPR_MODE_FAST_FP64;
    if (!std::strcmp(m, "plain")) return SUPR_MODE_PLAIN_FP64;
    if (!std::strcmp(m, "mixed")) return SUPR_MODE_MIXED;
    return SUPR_MODE_FAST_FP64;
}

inline int poll_stop(void* ctx) {
    return static_cast<std::atomic<bool>*>(ctx)->load(std::memory_order_relaxed) ? 1 : 0;
}

// solve_pdhg_gpu() (reference: include/pdlp_gpu.hpp:22-25) on the MI355X
// engine.  eta is ignored, as in the reference (the step size is adaptive).
template <class Result, class LP>
Result solve_pdhg_gpu_as(const LP& unscaled_form, const LP& scaled_form,
                         const std::vector<double>& Dr, const std::vector<double>& Dc,
                         double /*eta*/, int max_iterations, int check_every, double tol,
                         std::atomic<bool>* stop_flag, int mode = mode_from_env()) {
    const supr_lp U = view(unscaled_form), S = view(scaled_form);
    supr_pdlp_options o;
    supr_pdlp_default_options(&o);
    o.mode = mode;
    o.max_iterations = max_iterations;
    o.check_every = check_every;
    o.tol = tol;
    std::vector<double> x(S.n), y(S.m);
    supr_pdlp_info info;
    check(supr_solve_pdhg_gpu(&U, &S, Dr.data(), Dc.data(), &o, stop_flag ? poll_stop : nullptr,
                              stop_flag, x.data(), y.data(), &info),
          "supr_solve_pdhg_gpu");
    Result r;
    if (info.has_point) {
        r.x = std::move(x);
        r.y = std::move(y);
    }
    r.iterations = info.iterations;
    r.converged = info.converged != 0;
    r.eps_P = info.eps_P;
    r.eps_D = info.eps_D;
    r.eps_G = info.eps_G;
    r.restarts = info.restarts;
    return r;
}

}  // namespace suprema
