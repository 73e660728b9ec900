// drop_in_solve.cpp -- the C++ drop-in API (include/pdlp_gpu.hpp etc.) used
// the way the reference's callers use it (src/solve.cpp:71-148,
// tools/audit_lp.cpp:98-123): build the PDLP form, scale, solve on the GPU
// with a stop flag, extract, verify on the original LP.  Prints one JSON line;
// exit code 0 when every check passes.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <thread>
#include <vector>

#include "pdlp_gpu.hpp"
#include "verify.hpp"

static RangedLP transport(int S, int D) {
    RangedLP lp;
    const int n = S * D, m = S + D, nnz = 2 * n;
    lp.A.rows = m; lp.A.cols = n;
    lp.A.row_idx.resize(nnz); lp.A.col_idx.resize(nnz); lp.A.val.resize(nnz);
    lp.c.resize(n); lp.l.resize(n); lp.u.resize(n); lp.rL.resize(m); lp.rU.resize(m);
    suprema::check(supr_gen_transport(S, D, 42, 100, lp.A.row_idx.data(), lp.A.col_idx.data(),
                                      lp.A.val.data(), lp.c.data(), lp.rL.data(), lp.rU.data(),
                                      lp.l.data(), lp.u.data()), "gen");
    lp.col_names.resize(n); lp.row_names.resize(m);
    return lp;
}

int main() {
    int failures = 0;
    // 1. a full solve through the drop-in signature
    RangedLP lp = transport(60, 60);
    auto built = build_pdlp_form(lp);
    ScalingResult sr = scale(built.first);
    std::atomic<bool> stop{false};
    PdlpResult r = solve_pdhg_gpu(built.first, sr.scaled, sr.Dr, sr.Dc, 0.99, 200000, 64, 1e-6, &stop);
    std::vector<double> x, y;
    extract_solution(built.second, r.x, r.y, x, y);
    KKTReport k = verify(lp, x, y);
    const double worst = std::fmax(k.eps_P, std::fmax(k.eps_D, k.eps_G));
    failures += !(r.converged && worst <= 1e-6);
    // 2. cooperative cancellation from another thread (src/solve.cpp:96-99)
    RangedLP big = transport(300, 300);
    auto b2 = build_pdlp_form(big);
    ScalingResult s2 = scale(b2.first);
    std::atomic<bool> flag{false};
    std::thread setter([&] {
        std::this_thread::sleep_for(std::chrono::milliseconds(20));
        flag.store(true, std::memory_order_relaxed);
    });
    PdlpResult c = solve_pdhg_gpu(b2.first, s2.scaled, s2.Dr, s2.Dc, 0.99, 2000000, 64, 1e-12, &flag);
    setter.join();
    failures += c.converged ? 1 : 0;
    // 3. a device-side error surfaces as std::runtime_error
    bool threw = false;
    try {
        RangedLP bad = built.first;
        bad.A.row_idx[0] = bad.m() + 5;
        (void)solve_pdhg_gpu(bad, bad, sr.Dr, sr.Dc, 0.99, 100, 64, 1e-6, nullptr);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    failures += threw ? 0 : 1;
    std::printf("{\"converged\": %d, \"iterations\": %d, \"eps\": [%.3e, %.3e, %.3e], \"objective\": %.10g, "
                "\"cancelled_iterations\": %d, \"cancelled_converged\": %d, \"threw\": %d, \"failures\": %d}\n",
                (int)r.converged, r.iterations, k.eps_P, k.eps_D, k.eps_G, k.objective, c.iterations,
                (int)c.converged, (int)threw, failures);
    return failures == 0 ? 0 : 1;
}
