// server_tsan.cpp -- TEST PROGRAM (GPU): the serving engine with requests
// submitted from four threads at once while the main thread steps the
// server and collects results, built with ThreadSanitizer on the host code
// of the whole library (team-suprema_amd/Makefile target tsan_server).
// Every result must converge and pass the host verifier; prints "ok".
#include <atomic>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <thread>
#include <vector>

#include "suprema_pdlp.h"

struct Req {
    int32_t m, n, nnz;
    std::vector<int32_t> row, col;
    std::vector<double> val, c, rL, rU, l, u;                 // unscaled
    std::vector<double> vs, cs, rLs, rUs, ls, us, Dr, Dc;     // scaled
    supr_lp U() const { return {m, n, nnz, row.data(), col.data(), val.data(), c.data(), rL.data(), rU.data(), l.data(), u.data(), 0.0}; }
    supr_lp S() const { return {m, n, nnz, row.data(), col.data(), vs.data(), cs.data(), rLs.data(), rUs.data(), ls.data(), us.data(), 0.0}; }
};

static void make(Req& r, int m, int n, uint64_t seed) {
    r.m = m; r.n = n; r.nnz = 6 * n;
    r.row.resize(r.nnz); r.col.resize(r.nnz); r.val.resize(r.nnz);
    r.c.resize(n); r.l.resize(n); r.u.resize(n); r.rL.resize(m); r.rU.resize(m);
    supr_gen_random_lp(m, n, 6, nullptr, seed, 5.0, 10.0, 0, r.row.data(), r.col.data(), r.val.data(),
                       r.c.data(), r.rL.data(), r.rU.data(), r.l.data(), r.u.data());
    r.vs.resize(r.nnz); r.cs.resize(n); r.rLs.resize(m); r.rUs.resize(m); r.ls.resize(n); r.us.resize(n);
    r.Dr.resize(m); r.Dc.resize(n);
    const supr_lp u = r.U();
    supr_scale(&u, nullptr, 10, r.vs.data(), r.cs.data(), r.rLs.data(), r.rUs.data(), r.ls.data(),
               r.us.data(), r.Dr.data(), r.Dc.data());
}

int main() {
    if (supr_device_count() < 1) { std::puts("no device"); return 2; }
    supr_server_options so;
    supr_server_default_options(&so);
    so.max_cols = 1 << 20; so.max_rows = 1 << 19; so.max_nnz = 1 << 22; so.max_instances = 8;
    supr_server* srv = nullptr;
    if (supr_server_create(&so, &srv) != SUPR_OK) { std::printf("create: %s\n", supr_last_error()); return 1; }
    constexpr int T = 4, PER = 3;
    std::vector<Req> reqs(T * PER);
    std::vector<int64_t> tickets(T * PER, -1);
    std::mutex mu;
    std::atomic<int> failures{0};
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t]() {
            for (int k = 0; k < PER; ++k) {
                const int i = t * PER + k;
                make(reqs[i], 60 + 40 * i, 150 + 90 * i, 100 + i);
                const supr_lp U = reqs[i].U(), S = reqs[i].S();
                int64_t tk = -1;
                if (supr_server_submit(srv, &U, &S, reqs[i].Dr.data(), reqs[i].Dc.data(), &tk) != SUPR_OK) ++failures;
                std::lock_guard<std::mutex> lock(mu);
                tickets[i] = tk;
            }
        });
    std::vector<int> done(T * PER, 0);
    int ndone = 0, steps = 0;
    while (ndone < T * PER && steps < 2000000) {
        int32_t running, queued, finished;
        if (supr_server_step(srv, &running, &queued, &finished) != SUPR_OK) { ++failures; break; }
        ++steps;
        for (int i = 0; i < T * PER; ++i) {
            int64_t tk;
            {
                std::lock_guard<std::mutex> lock(mu);
                tk = tickets[i];
            }
            if (tk < 0 || done[i]) continue;
            int32_t m, n, ready = 0;
            if (supr_server_result(srv, tk, &m, &n, nullptr, nullptr, nullptr, &ready) != SUPR_OK) { ++failures; continue; }
            if (!ready) continue;
            std::vector<double> x(n), y(m);
            supr_pdlp_info info;
            supr_server_result(srv, tk, nullptr, nullptr, x.data(), y.data(), &info, &ready);
            double rep[7];
            const supr_lp U = reqs[i].U();
            supr_verify(&U, x.data(), y.data(), rep);
            const double worst = std::fmax(rep[0], std::fmax(rep[1], rep[2]));
            if (!info.converged || !(worst <= 1e-6)) ++failures;
            done[i] = 1;
            ++ndone;
        }
    }
    for (auto& t : th) t.join();
    supr_server_destroy(srv);
    std::printf("%d of %d requests verified in %d steps, %d failures\n%s\n", ndone, T * PER, steps,
                failures.load(), failures == 0 && ndone == T * PER ? "ok" : "FAILED");
    return failures == 0 && ndone == T * PER ? 0 : 1;
}
