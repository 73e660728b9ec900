// sanitize_host.cpp -- TEST PROGRAM: drives the library's host-side code
// (MPS reader / writer, generators, build_pdlp_form, scale, verify,
// extract_solution, argument checks) and the FP64 oracle through their C
// entry points, built without any device code under AddressSanitizer +
// UndefinedBehaviorSanitizer (or ThreadSanitizer: the threaded MPS reader).
//   sanitize_host <dir with .mps files> [threads]
// Exit status 0 and a last line "ok" when every check held; any sanitizer
// report aborts the program (halt_on_error).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dirent.h>
#include <string>
#include <vector>

#include "suprema_pdlp.h"

extern "C" {
int orc_verify(int m, int n, int nnz, const int* ri, const int* ci, const double* v, const double* c,
               const double* rL, const double* rU, const double* l, const double* u, double obj_offset,
               const double* x, const double* y, double* out7);
int orc_scale(int m, int n, int nnz, const int* ri, const int* ci, const double* v, const double* c,
              const double* rL, const double* rU, const double* l, const double* u, const char* integer,
              int ruiz_passes, double* val_s, double* c_s, double* rL_s, double* rU_s, double* l_s,
              double* u_s, double* Dr, double* Dc);
int orc_pdlp(int m, int n, int nnz, const int* ri, const int* ci, const double* val_u,
             const double* c_u, const double* rL_u, const double* rU_u, const double* l_u,
             const double* u_u, double obj_offset, const double* val_s, const double* c_s,
             const double* rL_s, const double* rU_s, const double* l_s, const double* u_s,
             const double* Dr, const double* Dc, int max_iterations, int check_every, double tol,
             const volatile int* stop, double* x_out, double* y_out, int* info_i, double* info_d);
}

static int failures = 0;
#define CHECK(c)                                                               \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::fprintf(stderr, "check failed at line %d: %s (%s)\n", __LINE__, #c, \
                         supr_last_error());                                   \
            ++failures;                                                        \
        }                                                                      \
    } while (0)

struct Lp {
    int32_t m = 0, n = 0, nnz = 0;
    std::vector<int32_t> row, col;
    std::vector<double> val, c, rL, rU, l, u;
    double off = 0.0;
    supr_lp view() const {
        return supr_lp{m, n, nnz, row.data(), col.data(), val.data(), c.data(), rL.data(), rU.data(),
                       l.data(), u.data(), off};
    }
};

// scale, PDLP form, verify, extract, and the oracle's scale / verify / a
// short PDLP run on one LP.
static void exercise(const Lp& lp, const char* what) {
    const supr_lp v = lp.view();
    int32_t dims[4];
    CHECK(supr_pdlp_form_sizes(&v, dims) == SUPR_OK);
    Lp f;
    f.m = dims[0]; f.n = dims[1]; f.nnz = dims[2]; f.off = lp.off;
    f.row.resize(f.nnz); f.col.resize(f.nnz); f.val.resize(f.nnz);
    f.c.resize(f.n); f.l.resize(f.n); f.u.resize(f.n); f.rL.resize(f.m); f.rU.resize(f.m);
    std::vector<int32_t> perm(f.m);
    std::vector<double> signs(f.m);
    CHECK(supr_pdlp_form(&v, f.row.data(), f.col.data(), f.val.data(), f.c.data(), f.rL.data(),
                         f.rU.data(), f.l.data(), f.u.data(), perm.data(), signs.data()) == SUPR_OK);
    const supr_lp fv = f.view();
    std::vector<double> vs(f.nnz), cs(f.n), rLs(f.m), rUs(f.m), ls(f.n), us(f.n), Dr(f.m), Dc(f.n);
    CHECK(supr_scale(&fv, nullptr, 10, vs.data(), cs.data(), rLs.data(), rUs.data(), ls.data(),
                     us.data(), Dr.data(), Dc.data()) == SUPR_OK);
    std::vector<double> vo(f.nnz), co(f.n), rLo(f.m), rUo(f.m), lo(f.n), uo(f.n), Dro(f.m), Dco(f.n);
    CHECK(orc_scale(f.m, f.n, f.nnz, f.row.data(), f.col.data(), f.val.data(), f.c.data(), f.rL.data(),
                    f.rU.data(), f.l.data(), f.u.data(), nullptr, 10, vo.data(), co.data(), rLo.data(),
                    rUo.data(), lo.data(), uo.data(), Dro.data(), Dco.data()) == 0);
    CHECK(std::memcmp(vs.data(), vo.data(), sizeof(double) * f.nnz) == 0);
    CHECK(std::memcmp(Dc.data(), Dco.data(), sizeof(double) * f.n) == 0);
    // a short solve by the oracle, then extract and verify on the LP
    std::vector<double> x2(f.n, 0.0), y2(f.m, 0.0), info_d(3);
    int info_i[4];
    CHECK(orc_pdlp(f.m, f.n, f.nnz, f.row.data(), f.col.data(), f.val.data(), f.c.data(), f.rL.data(),
                   f.rU.data(), f.l.data(), f.u.data(), f.off, vs.data(), cs.data(), rLs.data(),
                   rUs.data(), ls.data(), us.data(), Dr.data(), Dc.data(), 512, 64, 1e-6, nullptr,
                   x2.data(), y2.data(), info_i, info_d.data()) == 0);
    std::vector<double> x(lp.n), y(lp.m);
    CHECK(supr_extract_solution(lp.n, lp.m, f.m, perm.data(), signs.data(), x2.data(), y2.data(),
                                x.data(), y.data()) == SUPR_OK);
    double r1[7], r2[7];
    CHECK(supr_verify(&v, x.data(), y.data(), r1) == SUPR_OK);
    CHECK(orc_verify(lp.m, lp.n, lp.nnz, lp.row.data(), lp.col.data(), lp.val.data(), lp.c.data(),
                     lp.rL.data(), lp.rU.data(), lp.l.data(), lp.u.data(), lp.off, x.data(), y.data(),
                     r2) == 0);
    for (int q = 0; q < 7; ++q) CHECK(r1[q] == r2[q] || (std::isnan(r1[q]) && std::isnan(r2[q])));
    std::printf("%s: m=%d n=%d nnz=%d form %dx%d, verify eps %.2e %.2e %.2e\n", what, lp.m, lp.n,
                lp.nnz, f.m, f.n, r1[0], r1[1], r1[2]);
}

static Lp read(const std::string& path) {
    Lp lp;
    supr_mps* h = nullptr;
    if (supr_mps_read(path.c_str(), &h) != SUPR_OK) return lp;   // an error text is a valid outcome
    int32_t qn = 0, mx = 0;
    CHECK(supr_mps_sizes(h, &lp.m, &lp.n, &lp.nnz, &qn, &mx) == SUPR_OK);
    lp.row.resize(lp.nnz); lp.col.resize(lp.nnz); lp.val.resize(lp.nnz);
    lp.c.resize(lp.n); lp.l.resize(lp.n); lp.u.resize(lp.n); lp.rL.resize(lp.m); lp.rU.resize(lp.m);
    std::vector<char> integer(lp.n);
    std::vector<int32_t> qr(qn), qc(qn);
    std::vector<double> qv(qn);
    CHECK(supr_mps_get(h, lp.row.data(), lp.col.data(), lp.val.data(), lp.c.data(), lp.rL.data(),
                       lp.rU.data(), lp.l.data(), lp.u.data(), &lp.off, integer.data(), qr.data(),
                       qc.data(), qv.data()) == SUPR_OK);
    int64_t bytes = 0;
    CHECK(supr_mps_names(h, 1, nullptr, 0, &bytes) == SUPR_OK);
    std::vector<char> names((size_t)bytes + 1);
    CHECK(supr_mps_names(h, 1, names.data(), (int64_t)names.size(), &bytes) == SUPR_OK);
    CHECK(supr_mps_name(h, 2, lp.m) == nullptr);
    supr_mps_free(h);
    return lp;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <dir of .mps files> [threads]\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const bool threads_only = argc > 2;
    if (threads_only) {
        setenv("SUPR_MPS_THREADS", argv[2], 1);
        setenv("SUPR_MPS_MIN_CHUNK", "4096", 1);
    }
    // every fixture: read, and (LPs with rows and columns) the host pipeline
    DIR* d = opendir(dir.c_str());
    CHECK(d != nullptr);
    std::vector<std::string> files;
    while (d) {
        dirent* e = readdir(d);
        if (!e) break;
        const std::string nm = e->d_name;
        if (nm.size() > 4 && nm.substr(nm.size() - 4) == ".mps") files.push_back(dir + "/" + nm);
    }
    if (d) closedir(d);
    for (const auto& f : files) {
        const Lp lp = read(f);
        if (!threads_only && lp.m > 0 && lp.n > 0) exercise(lp, f.c_str());
    }
    // a generated transportation LP written as MPS and read back on several
    // threads (the threaded COLUMNS pass) must give the same arrays
    Lp t;
    const int S = 120, D = 150;
    t.m = S + D; t.n = S * D; t.nnz = 2 * S * D;
    t.row.resize(t.nnz); t.col.resize(t.nnz); t.val.resize(t.nnz);
    t.c.resize(t.n); t.l.resize(t.n); t.u.resize(t.n); t.rL.resize(t.m); t.rU.resize(t.m);
    CHECK(supr_gen_transport(S, D, 42, 100, t.row.data(), t.col.data(), t.val.data(), t.c.data(),
                             t.rL.data(), t.rU.data(), t.l.data(), t.u.data()) == SUPR_OK);
    const supr_lp tv = t.view();
    const std::string out = dir + "/roundtrip_sanitize.mps";
    CHECK(supr_mps_write(out.c_str(), &tv, "T", nullptr, nullptr, nullptr) == SUPR_OK);
    const Lp back = read(out);
    std::remove(out.c_str());
    CHECK(back.nnz == t.nnz && std::memcmp(back.val.data(), t.val.data(), sizeof(double) * t.nnz) == 0);
    CHECK(std::memcmp(back.c.data(), t.c.data(), sizeof(double) * t.n) == 0);
    if (!threads_only) {
        exercise(t, "transport 120x150");
        // random and blending generators, and argument errors
        Lp r;
        r.m = 60; r.n = 150; r.nnz = 150 * 6;
        r.row.resize(r.nnz); r.col.resize(r.nnz); r.val.resize(r.nnz);
        r.c.resize(r.n); r.l.resize(r.n); r.u.resize(r.n); r.rL.resize(r.m); r.rU.resize(r.m);
        CHECK(supr_gen_random_lp(r.m, r.n, 6, nullptr, 7, 5.0, 10.0, 0, r.row.data(), r.col.data(),
                                 r.val.data(), r.c.data(), r.rL.data(), r.rU.data(), r.l.data(),
                                 r.u.data()) == SUPR_OK);
        exercise(r, "random 60x150");
        Lp b;
        const int K = 12, P = 7;
        b.m = 3 * P + K; b.n = K * P; b.nnz = 4 * K * P;
        b.row.resize(b.nnz); b.col.resize(b.nnz); b.val.resize(b.nnz);
        b.c.resize(b.n); b.l.resize(b.n); b.u.resize(b.n); b.rL.resize(b.m); b.rU.resize(b.m);
        CHECK(supr_gen_blend(K, P, 3, b.row.data(), b.col.data(), b.val.data(), b.c.data(),
                             b.rL.data(), b.rU.data(), b.l.data(), b.u.data()) == SUPR_OK);
        exercise(b, "blend 12x7");
        Lp bad = r;
        bad.row[5] = bad.m;   // out of range
        const supr_lp bv = bad.view();
        double rep[7];
        std::vector<double> z(bad.n, 0.0), w(bad.m, 0.0);
        CHECK(supr_verify(&bv, z.data(), w.data(), rep) == SUPR_EINVAL);
        CHECK(std::strlen(supr_last_error()) > 0);
        supr_mps* h = nullptr;
        CHECK(supr_mps_read((dir + "/does_not_exist.mps").c_str(), &h) != SUPR_OK);
    }
    std::printf("%s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
