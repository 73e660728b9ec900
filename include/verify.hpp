// verify.hpp -- the independent FP64 KKT verifier (reference: include/
// verify.hpp:7-20, src/verify.cpp), and scale() (reference: include/
// scaling.hpp:7-13, src/scaling.cpp:25-84), over the C ABI.
#pragma once
#include <vector>

#include "ranged_lp.hpp"
#include "suprema_pdlp.hpp"

struct KKTReport {
    double eps_P = 0, eps_D = 0, eps_G = 0;
    double objective = 0, dual_objective = 0;
    double max_row_violation = 0, max_bound_violation = 0;
};

inline KKTReport verify(const RangedLP& ranged, const std::vector<double>& x,
                        const std::vector<double>& y) {
    const supr_lp v = suprema::view(ranged);
    double o[7];
    suprema::check(supr_verify(&v, x.data(), y.data(), o), "supr_verify");
    KKTReport r;
    r.eps_P = o[0]; r.eps_D = o[1]; r.eps_G = o[2];
    r.objective = o[3]; r.dual_objective = o[4];
    r.max_row_violation = o[5]; r.max_bound_violation = o[6];
    return r;
}

struct ScalingResult {
    RangedLP scaled;
    std::vector<double> Dr, Dc;
};

inline ScalingResult scale(const RangedLP& ranged, int ruiz_passes = 10) {
    const supr_lp v = suprema::view(ranged);
    ScalingResult s;
    s.scaled = ranged;
    s.scaled.name = ranged.name + " (scaled)";
    s.Dr.resize(ranged.m());
    s.Dc.resize(ranged.n());
    suprema::check(supr_scale(&v, nullptr, ruiz_passes, s.scaled.A.val.data(), s.scaled.c.data(),
                              s.scaled.rL.data(), s.scaled.rU.data(), s.scaled.l.data(),
                              s.scaled.u.data(), s.Dr.data(), s.Dc.data()),
                   "supr_scale");
    return s;
}
