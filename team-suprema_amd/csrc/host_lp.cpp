// host_lp.cpp -- host-side LP helpers of the engine and their C ABI:
// CSR layout, the FP64 KKT verifier (used by the plain mode's host check and
// by callers certifying a result), Ruiz + Pock-Chambolle scaling, the PDLP
// form, and the seeded LP generators.
//
// Reference behaviour followed (paths relative to the reference repository):
//   build_csr        include/sparse.hpp:31-57
//   verify_csr       src/verify.cpp:6-103
//   supr_scale       src/scaling.cpp:4-84 (LP part)
//   supr_pdlp_form   src/pdlp.cpp:8-82
//   supr_extract_solution  src/pdlp.cpp:84-90
//   supr_gen_transport     tools/gen_lp.cpp:63-130 (structure and value ranges)
#include "host_lp.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>

namespace supr {

namespace {
thread_local std::string t_error;
constexpr double kInf = std::numeric_limits<double>::infinity();
}  // namespace

void set_error(const std::string& msg) { t_error = msg; }

bool check_lp(const supr_lp* lp, const char* what, bool indices) {
    if (!lp) { set_error(std::string(what) + ": null LP"); return false; }
    if (lp->m < 0 || lp->n < 0 || lp->nnz < 0) {
        set_error(std::string(what) + ": negative size");
        return false;
    }
    if ((lp->nnz > 0 && (!lp->row_idx || !lp->col_idx || !lp->val)) ||
        (lp->n > 0 && (!lp->c || !lp->l || !lp->u)) || (lp->m > 0 && (!lp->rL || !lp->rU))) {
        set_error(std::string(what) + ": null array");
        return false;
    }
    if (!indices || lp->nnz == 0) return true;
    // one branch-free pass (unsigned compares fold the < 0 test in), then
    // the first bad entry only if there is one
    const uint32_t mu = (uint32_t)lp->m, nu = (uint32_t)lp->n;
    uint32_t bad = 0;
    for (int32_t k = 0; k < lp->nnz; ++k)
        bad |= (uint32_t)((uint32_t)lp->row_idx[k] >= mu) | (uint32_t)((uint32_t)lp->col_idx[k] >= nu);
    if (!bad) return true;
    for (int32_t k = 0; k < lp->nnz; ++k) {
        if (lp->row_idx[k] < 0 || lp->row_idx[k] >= lp->m || lp->col_idx[k] < 0 ||
            lp->col_idx[k] >= lp->n) {
            set_error(std::string(what) + ": COO index out of range at entry " + std::to_string(k));
            return false;
        }
    }
    return true;
}

Csr build_csr(const supr_lp& lp, const double* values, bool transposed) {
    Csr M;
    const int32_t* r = transposed ? lp.col_idx : lp.row_idx;
    const int32_t* c = transposed ? lp.row_idx : lp.col_idx;
    M.rows = transposed ? lp.n : lp.m;
    M.cols = transposed ? lp.m : lp.n;
    M.ptr.assign((size_t)M.rows + 1, 0);
    for (int32_t k = 0; k < lp.nnz; ++k) M.ptr[(size_t)r[k] + 1]++;
    for (int32_t i = 0; i < M.rows; ++i) M.ptr[(size_t)i + 1] += M.ptr[i];
    M.idx.resize(lp.nnz);
    M.val.resize(lp.nnz);
    M.perm.resize(lp.nnz);
    std::vector<int32_t> next(M.ptr.begin(), M.ptr.end() - 1);
    for (int32_t k = 0; k < lp.nnz; ++k) {
        const int32_t p = next[r[k]]++;
        M.idx[p] = c[k];
        M.val[p] = values ? values[k] : 0.0;
        M.perm[p] = k;
    }
    return M;
}

void csr_matvec(const Csr& M, const double* x, double* out) {
    for (int32_t i = 0; i < M.rows; ++i) {
        double s = 0.0;
        for (int32_t p = M.ptr[i]; p < M.ptr[(size_t)i + 1]; ++p) s += M.val[p] * x[M.idx[p]];
        out[i] = s;
    }
}

namespace {

double finite_inf_norm(const double* v, int32_t len) {
    double m = 0.0;
    bool any = false;
    for (int32_t k = 0; k < len; ++k)
        if (std::isfinite(v[k])) { m = std::max(m, std::fabs(v[k])); any = true; }
    return any ? m : 0.0;
}

}  // namespace

void verify_csr(const supr_lp& lp, const Csr& A, const Csr& AT, const double* x,
                const double* y, double* out7) {
    std::vector<double> ax(lp.m), aty(lp.n);
    csr_matvec(A, x, ax.data());
    csr_matvec(AT, y, aty.data());
    verify_products(lp, x, y, ax.data(), aty.data(), out7);
}

void verify_products(const supr_lp& lp, const double* x, const double* y, const double* ax,
                     const double* aty, double* out7) {
    const int32_t m = lp.m, n = lp.n;

    // Primal: row and bound violations, 2-norm over rows then columns.
    double sp = 0.0, max_rv = 0.0, max_bv = 0.0;
    for (int32_t i = 0; i < m; ++i) {
        const double v = std::max(lp.rL[i] - ax[i], 0.0) + std::max(ax[i] - lp.rU[i], 0.0);
        sp += v * v;
        if (i == 0 || v > max_rv) max_rv = v;
    }
    for (int32_t j = 0; j < n; ++j) {
        const double v = std::max(lp.l[j] - x[j], 0.0) + std::max(x[j] - lp.u[j], 0.0);
        sp += v * v;
        if (j == 0 || v > max_bv) max_bv = v;
    }
    const double denom_p =
        1.0 + std::max(finite_inf_norm(lp.rL, m), finite_inf_norm(lp.rU, m));

    // Dual: sign violations of y on infinite row bounds, of z = c - A'y on
    // infinite column bounds.
    double sd = 0.0;
    for (int32_t i = 0; i < m; ++i) {
        double v = 0.0;
        if (!std::isfinite(lp.rL[i]) && y[i] > 0) v += y[i];
        if (!std::isfinite(lp.rU[i]) && y[i] < 0) v += -y[i];
        sd += v * v;
    }
    std::vector<double> z(n);
    for (int32_t j = 0; j < n; ++j) {
        z[j] = lp.c[j] - aty[j];
        double v = 0.0;
        if (!std::isfinite(lp.l[j]) && z[j] > 0) v += z[j];
        if (!std::isfinite(lp.u[j]) && z[j] < 0) v += -z[j];
        sd += v * v;
    }
    const double denom_d = 1.0 + finite_inf_norm(lp.c, n);

    // Dual objective: multipliers on infinite bounds are dropped (they are
    // already counted as sign violations), so D stays finite.
    double D = 0.0;
    for (int32_t i = 0; i < m; ++i) {
        double yv = y[i];
        if (!std::isfinite(lp.rL[i]) && yv > 0) yv = 0.0;
        if (!std::isfinite(lp.rU[i]) && yv < 0) yv = 0.0;
        if (yv > 0 && std::isfinite(lp.rL[i])) D += lp.rL[i] * yv;
        if (yv < 0 && std::isfinite(lp.rU[i])) D += -lp.rU[i] * -yv;
    }
    for (int32_t j = 0; j < n; ++j) {
        double zv = z[j];
        if (!std::isfinite(lp.l[j]) && zv > 0) zv = 0.0;
        if (!std::isfinite(lp.u[j]) && zv < 0) zv = 0.0;
        if (zv > 0 && std::isfinite(lp.l[j])) D += lp.l[j] * zv;
        if (zv < 0 && std::isfinite(lp.u[j])) D += -lp.u[j] * -zv;
    }
    double cx = 0.0;
    for (int32_t j = 0; j < n; ++j) cx += lp.c[j] * x[j];

    out7[0] = std::sqrt(sp) / denom_p;
    out7[1] = std::sqrt(sd) / denom_d;
    out7[2] = std::isfinite(D) ? std::fabs(cx - D) / (1.0 + std::fabs(cx) + std::fabs(D)) : kInf;
    out7[3] = cx + lp.obj_offset;
    out7[4] = std::isfinite(D) ? D + lp.obj_offset : D;
    out7[5] = m > 0 ? max_rv : 0.0;
    out7[6] = n > 0 ? max_bv : 0.0;
}

}  // namespace supr

using supr::set_error;

extern "C" {

const char* supr_last_error(void) { return supr::t_error.c_str(); }

const char* supr_version(void) { return "suprema-pdlp-gfx950 0.1"; }

int supr_verify(const supr_lp* lp, const double* x, const double* y, double* out7) {
    if (!supr::check_lp(lp, "supr_verify")) return SUPR_EINVAL;
    if ((lp->n > 0 && !x) || (lp->m > 0 && !y) || !out7) {
        set_error("supr_verify: null x, y or out");
        return SUPR_EINVAL;
    }
    try {
        supr::Csr A = supr::build_csr(*lp, lp->val, false);
        supr::Csr AT = supr::build_csr(*lp, lp->val, true);
        supr::verify_csr(*lp, A, AT, x, y, out7);
        return SUPR_OK;
    } catch (const std::bad_alloc&) {
        set_error("supr_verify: out of host memory");
        return SUPR_ENOMEM;
    }
}

// Ruiz equilibration (row then column inf-norms, 1/sqrt) followed by the
// Pock-Chambolle alpha = 1 diagonal step; the arithmetic order is the
// reference's, so the scaled LP is bit-identical to scale()'s.
int supr_scale(const supr_lp* lp, const char* integer, int32_t ruiz_passes, double* val_s,
               double* c_s, double* rL_s, double* rU_s, double* l_s, double* u_s, double* Dr,
               double* Dc) {
    if (!supr::check_lp(lp, "supr_scale")) return SUPR_EINVAL;
    if (ruiz_passes < 0) { set_error("supr_scale: negative ruiz_passes"); return SUPR_EINVAL; }
    try {
        const int32_t m = lp->m, n = lp->n, nnz = lp->nnz;
        const int32_t* ri = lp->row_idx;
        const int32_t* ci = lp->col_idx;
        std::vector<double> a(lp->val, lp->val + nnz), rs(m), cs(n);
        std::fill(Dr, Dr + m, 1.0);
        std::fill(Dc, Dc + n, 1.0);
        auto keep = [&](int32_t j) { return integer != nullptr && integer[j] != 0; };
        for (int32_t pass = 0; pass < ruiz_passes; ++pass) {
            std::fill(rs.begin(), rs.end(), 0.0);
            for (int32_t k = 0; k < nnz; ++k) rs[ri[k]] = std::max(rs[ri[k]], std::fabs(a[k]));
            for (int32_t i = 0; i < m; ++i) rs[i] = rs[i] > 0 ? 1.0 / std::sqrt(rs[i]) : 1.0;
            for (int32_t k = 0; k < nnz; ++k) a[k] = a[k] * rs[ri[k]];
            for (int32_t i = 0; i < m; ++i) Dr[i] *= rs[i];
            std::fill(cs.begin(), cs.end(), 0.0);
            for (int32_t k = 0; k < nnz; ++k) cs[ci[k]] = std::max(cs[ci[k]], std::fabs(a[k]));
            for (int32_t j = 0; j < n; ++j) cs[j] = (cs[j] > 0 && !keep(j)) ? 1.0 / std::sqrt(cs[j]) : 1.0;
            for (int32_t k = 0; k < nnz; ++k) a[k] = a[k] * cs[ci[k]];
            for (int32_t j = 0; j < n; ++j) Dc[j] *= cs[j];
        }
        std::fill(rs.begin(), rs.end(), 0.0);
        std::fill(cs.begin(), cs.end(), 0.0);
        for (int32_t k = 0; k < nnz; ++k) rs[ri[k]] += std::fabs(a[k]);
        for (int32_t k = 0; k < nnz; ++k) cs[ci[k]] += std::fabs(a[k]);
        for (int32_t i = 0; i < m; ++i) rs[i] = std::sqrt(rs[i] > 0 ? 1.0 / rs[i] : 1.0);
        for (int32_t j = 0; j < n; ++j) cs[j] = keep(j) ? 1.0 : std::sqrt(cs[j] > 0 ? 1.0 / cs[j] : 1.0);
        for (int32_t k = 0; k < nnz; ++k) val_s[k] = a[k] * rs[ri[k]] * cs[ci[k]];
        for (int32_t i = 0; i < m; ++i) Dr[i] *= rs[i];
        for (int32_t j = 0; j < n; ++j) Dc[j] *= cs[j];
        for (int32_t j = 0; j < n; ++j) {
            c_s[j] = Dc[j] * lp->c[j];
            l_s[j] = lp->l[j] / Dc[j];
            u_s[j] = lp->u[j] / Dc[j];
        }
        for (int32_t i = 0; i < m; ++i) {
            rL_s[i] = Dr[i] * lp->rL[i];
            rU_s[i] = Dr[i] * lp->rU[i];
        }
        return SUPR_OK;
    } catch (const std::bad_alloc&) {
        set_error("supr_scale: out of host memory");
        return SUPR_ENOMEM;
    }
}

namespace {
// Row classes of the PDLP form: equality, ranged (equality row + slack
// column), <= (negated into >=), >=, free (dropped).
enum RowKind { kEq, kRanged, kLe, kGe, kFree };
RowKind row_kind(double lo, double hi) {
    if (lo == hi) return kEq;
    if (std::isfinite(lo) && std::isfinite(hi)) return kRanged;
    if (std::isfinite(hi)) return kLe;
    if (std::isfinite(lo)) return kGe;
    return kFree;
}
}  // namespace

int supr_pdlp_form_sizes(const supr_lp* lp, int32_t* dims4) {
    if (!supr::check_lp(lp, "supr_pdlp_form_sizes") || !dims4) return SUPR_EINVAL;
    int32_t eq = 0, ineq = 0, slack = 0, nnz2 = 0;
    for (int32_t i = 0; i < lp->m; ++i) {
        const RowKind k = row_kind(lp->rL[i], lp->rU[i]);
        eq += (k == kEq || k == kRanged);
        slack += (k == kRanged);
        ineq += (k == kLe || k == kGe);
    }
    for (int32_t k = 0; k < lp->nnz; ++k)
        nnz2 += row_kind(lp->rL[lp->row_idx[k]], lp->rU[lp->row_idx[k]]) != kFree;
    dims4[0] = eq + ineq;
    dims4[1] = lp->n + slack;
    dims4[2] = nnz2 + slack;
    dims4[3] = slack;
    return SUPR_OK;
}

// Rows are reordered equality-first (original order within each class);
// <= rows are negated into >= rows; a ranged row becomes an equality row
// with b = 0 and a slack column (coefficient -1) carrying the range.
int supr_pdlp_form(const supr_lp* lp, int32_t* row_idx2, int32_t* col_idx2, double* val2,
                   double* c2, double* rL2, double* rU2, double* l2, double* u2,
                   int32_t* row_perm, double* row_signs) {
    if (!supr::check_lp(lp, "supr_pdlp_form")) return SUPR_EINVAL;
    const int32_t m = lp->m, n = lp->n;
    std::vector<int32_t> pos_of(m, -1), slack_of;
    std::vector<double> slo, shi;
    int32_t p = 0;
    for (int32_t i = 0; i < m; ++i) {
        const RowKind k = row_kind(lp->rL[i], lp->rU[i]);
        if (k != kEq && k != kRanged) continue;
        row_perm[p] = i;
        row_signs[p] = 1.0;
        if (k == kEq) {
            rL2[p] = rU2[p] = lp->rL[i];
            slack_of.push_back(-1);
        } else {
            rL2[p] = rU2[p] = 0.0;
            slack_of.push_back((int32_t)slo.size());
            slo.push_back(lp->rL[i]);
            shi.push_back(lp->rU[i]);
        }
        pos_of[i] = p++;
    }
    const int32_t n_eq = p;
    for (int32_t i = 0; i < m; ++i) {
        const RowKind k = row_kind(lp->rL[i], lp->rU[i]);
        if (k != kLe && k != kGe) continue;
        row_perm[p] = i;
        row_signs[p] = k == kLe ? -1.0 : 1.0;
        rL2[p] = k == kLe ? -lp->rU[i] : lp->rL[i];
        rU2[p] = supr::kInf;
        pos_of[i] = p++;
    }
    int32_t q = 0;
    for (int32_t k = 0; k < lp->nnz; ++k) {
        const int32_t pos = pos_of[lp->row_idx[k]];
        if (pos < 0) continue;
        row_idx2[q] = pos;
        col_idx2[q] = lp->col_idx[k];
        val2[q] = row_signs[pos] * lp->val[k];
        ++q;
    }
    for (int32_t pos = 0; pos < n_eq; ++pos) {
        if (slack_of[pos] < 0) continue;
        row_idx2[q] = pos;
        col_idx2[q] = n + slack_of[pos];
        val2[q] = -1.0;
        ++q;
    }
    std::memcpy(c2, lp->c, sizeof(double) * n);
    std::memcpy(l2, lp->l, sizeof(double) * n);
    std::memcpy(u2, lp->u, sizeof(double) * n);
    for (size_t s = 0; s < slo.size(); ++s) {
        c2[n + s] = 0.0;
        l2[n + s] = slo[s];
        u2[n + s] = shi[s];
    }
    return SUPR_OK;
}

int supr_extract_solution(int32_t n_orig, int32_t m_orig, int32_t m2, const int32_t* row_perm,
                          const double* row_signs, const double* x2, const double* y2,
                          double* x, double* y) {
    if (n_orig < 0 || m_orig < 0 || m2 < 0) { set_error("supr_extract_solution: negative size"); return SUPR_EINVAL; }
    std::memcpy(x, x2, sizeof(double) * n_orig);
    std::fill(y, y + m_orig, 0.0);
    for (int32_t pos = 0; pos < m2; ++pos) {
        if (row_perm[pos] < 0 || row_perm[pos] >= m_orig) { set_error("supr_extract_solution: bad row_perm"); return SUPR_EINVAL; }
        y[row_perm[pos]] = row_signs[pos] * y2[pos];
    }
    return SUPR_OK;
}

}  // extern "C"
