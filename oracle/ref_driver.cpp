// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C-ABI shim around the reference's own CPU PDLP sources, compiled from the
// reference tree by oracle/Makefile into oracle/_ref/libsovref.so.  The entry
// points take the same raw arrays as the oracle's (orc_*) functions, build
// the reference's RangedLP from them and call the reference functions
// unchanged:
//   ref_scale        -> scale()              src/scaling.cpp:25-84
//   ref_verify       -> verify()             src/verify.cpp:18-103
//   ref_pdlp_form_*  -> build_pdlp_form()    src/pdlp.cpp:8-82
//   ref_extract_solution -> extract_solution() src/pdlp.cpp:84-90
//   ref_pdlp[_live]  -> solve_pdhg()         src/pdlp.cpp:146-152 (pdlp_run)
//   ref_mps_*        -> read_mps() + to_ranged_lp() (arrays, integer flags,
//                       QP triangle, names, error text)
//                       src/mps_reader.cpp:54-308, include/ranged_lp.hpp:39-62
// It exists to pin the oracle and to time the reference CPU engine; the
// product never loads it.
#include "mps_reader.hpp"
#include "pdlp.hpp"
#include "ranged_lp.hpp"
#include "scaling.hpp"
#include "verify.hpp"

#include <atomic>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

namespace {

RangedLP make_lp(int m, int n, int nnz, const int* ri, const int* ci, const double* v,
                 const double* c, const double* rL, const double* rU, const double* l,
                 const double* u, double obj_offset) {
    RangedLP r;
    r.name = "lp";
    r.col_names.resize(n);
    r.row_names.resize(m);
    for (int j = 0; j < n; ++j) r.col_names[j] = "c" + std::to_string(j);
    for (int i = 0; i < m; ++i) r.row_names[i] = "r" + std::to_string(i);
    r.c.assign(c, c + n);
    r.A.rows = m; r.A.cols = n;
    r.A.row_idx.assign(ri, ri + nnz);
    r.A.col_idx.assign(ci, ci + nnz);
    r.A.val.assign(v, v + nnz);
    r.rL.assign(rL, rL + m); r.rU.assign(rU, rU + m);
    r.l.assign(l, l + n); r.u.assign(u, u + n);
    r.obj_offset = obj_offset;
    r.original_sense = "min";
    return r;
}

RangedLP g_mps;   // last MPS read (single-threaded test use)

}  // namespace

extern "C" {

int ref_scale(int m, int n, int nnz, const int* ri, const int* ci, const double* v,
              const double* c, const double* rL, const double* rU, const double* l,
              const double* u, int ruiz_passes, double* val_s, double* c_s, double* rL_s,
              double* rU_s, double* l_s, double* u_s, double* Dr, double* Dc) {
    try {
        RangedLP r = make_lp(m, n, nnz, ri, ci, v, c, rL, rU, l, u, 0.0);
        ScalingResult s = scale(r, ruiz_passes);
        std::memcpy(val_s, s.scaled.A.val.data(), sizeof(double) * nnz);
        std::memcpy(c_s, s.scaled.c.data(), sizeof(double) * n);
        std::memcpy(rL_s, s.scaled.rL.data(), sizeof(double) * m);
        std::memcpy(rU_s, s.scaled.rU.data(), sizeof(double) * m);
        std::memcpy(l_s, s.scaled.l.data(), sizeof(double) * n);
        std::memcpy(u_s, s.scaled.u.data(), sizeof(double) * n);
        std::memcpy(Dr, s.Dr.data(), sizeof(double) * m);
        std::memcpy(Dc, s.Dc.data(), sizeof(double) * n);
        return 0;
    } catch (const std::exception&) { return -1; }
}

int ref_verify(int m, int n, int nnz, const int* ri, const int* ci, const double* v,
               const double* c, const double* rL, const double* rU, const double* l,
               const double* u, double obj_offset, const double* x, const double* y,
               double* out7) {
    try {
        RangedLP r = make_lp(m, n, nnz, ri, ci, v, c, rL, rU, l, u, obj_offset);
        std::vector<double> xv(x, x + n), yv(y, y + m);
        KKTReport k = verify(r, xv, yv);
        out7[0] = k.eps_P; out7[1] = k.eps_D; out7[2] = k.eps_G;
        out7[3] = k.objective; out7[4] = k.dual_objective;
        out7[5] = k.max_row_violation; out7[6] = k.max_bound_violation;
        return 0;
    } catch (const std::exception&) { return -1; }
}

int ref_pdlp_form_sizes(int m, int n, int nnz, const int* ri, const int* ci, const double* v,
                        const double* c, const double* rL, const double* rU, const double* l,
                        const double* u, int* dims) {
    try {
        RangedLP r = make_lp(m, n, nnz, ri, ci, v, c, rL, rU, l, u, 0.0);
        auto f = build_pdlp_form(r);
        dims[0] = f.first.m(); dims[1] = f.first.n(); dims[2] = f.first.A.nnz();
        dims[3] = f.second.n_slack;
        return 0;
    } catch (const std::exception&) { return -1; }
}

int ref_pdlp_form(int m, int n, int nnz, const int* ri, const int* ci, const double* v,
                  const double* c, const double* rL, const double* rU, const double* l,
                  const double* u, int* ri2, int* ci2, double* v2, double* c2, double* rL2,
                  double* rU2, double* l2, double* u2, int* row_perm, double* row_signs) {
    try {
        RangedLP r = make_lp(m, n, nnz, ri, ci, v, c, rL, rU, l, u, 0.0);
        auto f = build_pdlp_form(r);
        const RangedLP& p = f.first;
        const int nz = p.A.nnz(), m2 = p.m(), n2 = p.n();
        std::memcpy(ri2, p.A.row_idx.data(), sizeof(int) * nz);
        std::memcpy(ci2, p.A.col_idx.data(), sizeof(int) * nz);
        std::memcpy(v2, p.A.val.data(), sizeof(double) * nz);
        std::memcpy(c2, p.c.data(), sizeof(double) * n2);
        std::memcpy(rL2, p.rL.data(), sizeof(double) * m2);
        std::memcpy(rU2, p.rU.data(), sizeof(double) * m2);
        std::memcpy(l2, p.l.data(), sizeof(double) * n2);
        std::memcpy(u2, p.u.data(), sizeof(double) * n2);
        std::memcpy(row_perm, f.second.row_perm.data(), sizeof(int) * m2);
        std::memcpy(row_signs, f.second.row_signs.data(), sizeof(double) * m2);
        return 0;
    } catch (const std::exception&) { return -1; }
}

// Same argument list as the oracle's orc_extract_solution: meta is rebuilt
// from its three fields (extract_solution reads n_orig, m_orig, row_perm and
// row_signs only).
int ref_extract_solution(int n_orig, int m_orig, int m2, const int* row_perm,
                         const double* row_signs, const double* x2, const double* y2,
                         double* x, double* y) {
    try {
        PdlpFormMeta meta;
        meta.n_orig = n_orig;
        meta.m_orig = m_orig;
        meta.row_perm.assign(row_perm, row_perm + m2);
        meta.row_signs.assign(row_signs, row_signs + m2);
        int n2 = n_orig;
        // x2 holds the form's n2 >= n_orig columns; only the first n_orig are read
        std::vector<double> vx2(x2, x2 + n2), vy2(y2, y2 + m2), vx, vy;
        extract_solution(meta, vx2, vy2, vx, vy);
        if ((int)vx.size() != n_orig || (int)vy.size() != m_orig) return -1;
        std::memcpy(x, vx.data(), sizeof(double) * vx.size());
        std::memcpy(y, vy.data(), sizeof(double) * vy.size());
        return 0;
    } catch (const std::exception&) { return -1; }
}

// Same argument list as the oracle's orc_pdlp.
int ref_pdlp(int m, int n, int nnz, const int* ri, const int* ci,
             const double* val_u, const double* c_u, const double* rL_u, const double* rU_u,
             const double* l_u, const double* u_u, double obj_offset,
             const double* val_s, const double* c_s, const double* rL_s, const double* rU_s,
             const double* l_s, const double* u_s, const double* Dr, const double* Dc,
             int max_iterations, int check_every, double tol, const volatile int* stop,
             double* x_out, double* y_out, int* info_i, double* info_d) {
    try {
        RangedLP U = make_lp(m, n, nnz, ri, ci, val_u, c_u, rL_u, rU_u, l_u, u_u, obj_offset);
        RangedLP S = make_lp(m, n, nnz, ri, ci, val_s, c_s, rL_s, rU_s, l_s, u_s, obj_offset);
        std::vector<double> dr(Dr, Dr + m), dc(Dc, Dc + n);
        std::atomic<bool> flag{stop && *stop};
        PdlpResult r = solve_pdhg(U, S, dr, dc, 0.99, max_iterations, check_every, tol,
                                  stop ? &flag : nullptr);
        const bool has = r.x.size() == (size_t)n && r.y.size() == (size_t)m;
        if (has) {
            std::memcpy(x_out, r.x.data(), sizeof(double) * n);
            std::memcpy(y_out, r.y.data(), sizeof(double) * m);
        }
        info_i[0] = r.iterations; info_i[1] = r.restarts; info_i[2] = r.converged ? 1 : 0;
        info_i[3] = has ? 1 : 0;
        info_d[0] = r.eps_P; info_d[1] = r.eps_D; info_d[2] = r.eps_G;
        return 0;
    } catch (const std::exception&) { return -1; }
}

// read_mps + to_ranged_lp; dims[0..2] = m, n, nnz.  The arrays are fetched
// with ref_mps_arrays (the LP is held until the next ref_mps_read).
int ref_mps_read(const char* path, int* dims) {
    try {
        g_mps = to_ranged_lp(read_mps(path));
        dims[0] = g_mps.m(); dims[1] = g_mps.n(); dims[2] = g_mps.A.nnz();
        return 0;
    } catch (const std::exception&) { return -1; }
}

int ref_mps_arrays(int* ri, int* ci, double* v, double* c, double* rL, double* rU, double* l,
                   double* u, double* obj_offset) {
    const RangedLP& r = g_mps;
    const int nz = r.A.nnz(), m = r.m(), n = r.n();
    std::memcpy(ri, r.A.row_idx.data(), sizeof(int) * nz);
    std::memcpy(ci, r.A.col_idx.data(), sizeof(int) * nz);
    std::memcpy(v, r.A.val.data(), sizeof(double) * nz);
    std::memcpy(c, r.c.data(), sizeof(double) * n);
    std::memcpy(rL, r.rL.data(), sizeof(double) * m);
    std::memcpy(rU, r.rU.data(), sizeof(double) * m);
    std::memcpy(l, r.l.data(), sizeof(double) * n);
    std::memcpy(u, r.u.data(), sizeof(double) * n);
    *obj_offset = r.obj_offset;
    return r.original_sense == "max" ? 1 : 0;
}

// ref_pdlp with a live stop flag: flag points at a std::atomic<bool> the
// caller may set from another thread while solve_pdhg runs (its relaxed load
// once per outer iteration, include/pdlp_algo.hpp:104) -- for race tests.
int ref_pdlp_live(int m, int n, int nnz, const int* ri, const int* ci,
                  const double* val_u, const double* c_u, const double* rL_u, const double* rU_u,
                  const double* l_u, const double* u_u, double obj_offset,
                  const double* val_s, const double* c_s, const double* rL_s, const double* rU_s,
                  const double* l_s, const double* u_s, const double* Dr, const double* Dc,
                  int max_iterations, int check_every, double tol, void* flag,
                  double* x_out, double* y_out, int* info_i, double* info_d) {
    try {
        RangedLP U = make_lp(m, n, nnz, ri, ci, val_u, c_u, rL_u, rU_u, l_u, u_u, obj_offset);
        RangedLP S = make_lp(m, n, nnz, ri, ci, val_s, c_s, rL_s, rU_s, l_s, u_s, obj_offset);
        std::vector<double> dr(Dr, Dr + m), dc(Dc, Dc + n);
        PdlpResult r = solve_pdhg(U, S, dr, dc, 0.99, max_iterations, check_every, tol,
                                  static_cast<std::atomic<bool>*>(flag));
        const bool has = r.x.size() == (size_t)n && r.y.size() == (size_t)m;
        if (has) {
            std::memcpy(x_out, r.x.data(), sizeof(double) * n);
            std::memcpy(y_out, r.y.data(), sizeof(double) * m);
        }
        info_i[0] = r.iterations; info_i[1] = r.restarts; info_i[2] = r.converged ? 1 : 0;
        info_i[3] = has ? 1 : 0;
        info_d[0] = r.eps_P; info_d[1] = r.eps_D; info_d[2] = r.eps_G;
        return 0;
    } catch (const std::exception&) { return -1; }
}

// The rest of the last MPS read: integer flags (n bytes), the QP lower
// triangle (ref_mps_q_nnz entries), and names (kind 0 problem, 1 column,
// 2 row; NULL when out of range).
int ref_mps_q_nnz() { return g_mps.Q.nnz(); }

int ref_mps_extra(char* integer, int* q_row, int* q_col, double* q_val) {
    const RangedLP& r = g_mps;
    for (int j = 0; j < r.n(); ++j) integer[j] = (j < (int)r.integer.size()) ? r.integer[j] : 0;
    const int qn = r.Q.nnz();
    std::memcpy(q_row, r.Q.row_idx.data(), sizeof(int) * qn);
    std::memcpy(q_col, r.Q.col_idx.data(), sizeof(int) * qn);
    std::memcpy(q_val, r.Q.val.data(), sizeof(double) * qn);
    return 0;
}

const char* ref_mps_name(int kind, int index) {
    const RangedLP& r = g_mps;
    if (kind == 0) return r.name.c_str();
    if (kind == 1) return index >= 0 && index < r.n() ? r.col_names[index].c_str() : nullptr;
    if (kind == 2) return index >= 0 && index < r.m() ? r.row_names[index].c_str() : nullptr;
    return nullptr;
}

// What the reference's read_mps throws on a file (its what() text), or an
// empty string when it parses.
const char* ref_mps_error(const char* path) {
    static thread_local std::string msg;
    try {
        (void)read_mps(path);
        msg.clear();
    } catch (const std::exception& e) {
        msg = e.what();
    }
    return msg.c_str();
}

}  // extern "C"
