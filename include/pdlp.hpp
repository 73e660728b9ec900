// pdlp.hpp -- the PDLP form and result types of the path (reference:
// include/pdlp.hpp:11-28), over the MI355X engine's C ABI.  build_pdlp_form /
// extract_solution run on the host, like the reference's.
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "ranged_lp.hpp"
#include "suprema_pdlp.hpp"

struct PdlpFormMeta {
    int n_orig = 0, m_orig = 0, n_slack = 0;
    std::vector<int> row_perm;
    std::vector<double> row_signs;
    std::vector<int> free_rows;
};

struct PdlpResult {
    std::vector<double> x, y;
    int iterations = 0;
    bool converged = false;
    double eps_P = 0, eps_D = 0, eps_G = 0;
    int restarts = 0;
};

// Equality (and ranged) rows first, then <= rows negated into >= rows; a
// ranged row gets a slack column; free rows are dropped (reference:
// src/pdlp.cpp:8-82).
inline std::pair<RangedLP, PdlpFormMeta> build_pdlp_form(const RangedLP& ranged) {
    const supr_lp v = suprema::view(ranged);
    int32_t d[4];
    suprema::check(supr_pdlp_form_sizes(&v, d), "supr_pdlp_form_sizes");
    RangedLP f;
    f.name = ranged.name + " (pdlp-form)";
    f.A.rows = d[0];
    f.A.cols = d[1];
    f.A.row_idx.resize(d[2]);
    f.A.col_idx.resize(d[2]);
    f.A.val.resize(d[2]);
    f.c.resize(d[1]); f.l.resize(d[1]); f.u.resize(d[1]);
    f.rL.resize(d[0]); f.rU.resize(d[0]);
    PdlpFormMeta meta;
    meta.n_orig = ranged.n();
    meta.m_orig = ranged.m();
    meta.n_slack = d[3];
    meta.row_perm.resize(d[0]);
    meta.row_signs.resize(d[0]);
    suprema::check(supr_pdlp_form(&v, f.A.row_idx.data(), f.A.col_idx.data(), f.A.val.data(),
                                  f.c.data(), f.rL.data(), f.rU.data(), f.l.data(), f.u.data(),
                                  meta.row_perm.data(), meta.row_signs.data()),
                   "supr_pdlp_form");
    f.col_names = ranged.col_names;
    f.col_names.resize(d[1]);
    for (int k = 0; k < d[3]; ++k) f.col_names[ranged.n() + k] = "__range_slack_" + std::to_string(k);
    f.row_names.resize(d[0]);
    for (int p = 0; p < d[0]; ++p)
        f.row_names[p] = meta.row_perm[p] < (int)ranged.row_names.size()
                             ? ranged.row_names[meta.row_perm[p]] : std::string();
    f.obj_offset = ranged.obj_offset;
    f.original_sense = ranged.original_sense;
    for (int i = 0; i < ranged.m(); ++i)
        if (!(ranged.rL[i] > -INF) && !(ranged.rU[i] < INF)) meta.free_rows.push_back(i);
    return {f, meta};
}

inline void extract_solution(const PdlpFormMeta& meta, const std::vector<double>& x2,
                             const std::vector<double>& y2, std::vector<double>& x,
                             std::vector<double>& y) {
    x.resize(meta.n_orig);
    y.resize(meta.m_orig);
    suprema::check(supr_extract_solution(meta.n_orig, meta.m_orig, (int32_t)meta.row_perm.size(),
                                         meta.row_perm.data(), meta.row_signs.data(), x2.data(),
                                         y2.data(), x.data(), y.data()),
                   "supr_extract_solution");
}
