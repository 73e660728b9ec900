// race.cpp -- the reference's CPU / GPU race (src/solve.cpp:76-148, the
// crossover race :290 and portfolio engine 3 :560 follow the same pattern)
// with this repository's GPU engine behind the drop-in signature and the
// reference's own CPU engine (solve_pdhg, src/pdlp.cpp:146-152, compiled into
// oracle/_ref/libsovref.so and loaded at run time) on the other thread.
//
// The race itself is written here from the reference's stated contract: both
// engines get the same const inputs and one std::atomic<bool> stop flag; the
// first to return a converged answer wins a compare-exchange and sets the
// flag; the other returns at its next poll; an engine's exception is caught on
// its own thread.  Checks, printed as one JSON line (exit 0 when all hold):
//   1. on a 300,000-variable transportation LP the GPU wins, its answer
//      verifies on the original LP, and the CPU engine is stopped early;
//   2. with the GPU started 300 ms late (a staggered start) on a small LP the
//      CPU wins, and the GPU engine, finding the flag set, returns within
//      one poll of its loop without converging.
// Usage: race <path to libsovref.so>
#include <dlfcn.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "pdlp_gpu.hpp"
#include "verify.hpp"

namespace {

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

using RefPdlp = int (*)(int, int, int, const int*, const int*, const double*, const double*,
                        const double*, const double*, const double*, const double*, double,
                        const double*, const double*, const double*, const double*, const double*,
                        const double*, const double*, const double*, int, int, double, void*,
                        double*, double*, int*, double*);
RefPdlp g_ref = nullptr;

// The reference's solve_pdhg on one thread, through the shim.
PdlpResult cpu_solve(const RangedLP& U, const RangedLP& S, const std::vector<double>& Dr,
                     const std::vector<double>& Dc, int max_it, double tol, std::atomic<bool>* flag) {
    PdlpResult r;
    std::vector<double> x(U.n()), y(U.m());
    int ii[4] = {0, 0, 0, 0};
    double dd[3] = {0, 0, 0};
    const int rc = g_ref(U.m(), U.n(), U.A.nnz(), U.A.row_idx.data(), U.A.col_idx.data(), U.A.val.data(),
                         U.c.data(), U.rL.data(), U.rU.data(), U.l.data(), U.u.data(), U.obj_offset,
                         S.A.val.data(), S.c.data(), S.rL.data(), S.rU.data(), S.l.data(), S.u.data(),
                         Dr.data(), Dc.data(), max_it, 64, tol, flag, x.data(), y.data(), ii, dd);
    if (rc != 0) throw std::runtime_error("reference CPU engine failed");
    if (ii[3]) { r.x = x; r.y = y; }
    r.iterations = ii[0];
    r.restarts = ii[1];
    r.converged = ii[2] != 0;
    r.eps_P = dd[0]; r.eps_D = dd[1]; r.eps_G = dd[2];
    return r;
}

struct Race {
    int winner = -1;   // 0 cpu, 1 gpu
    PdlpResult cpu, gpu;
    bool cpu_ok = false, gpu_ok = false;
    std::string cpu_err, gpu_err;
    double cpu_s = 0, gpu_s = 0;   // each engine's own wall time
};

Race race(const RangedLP& U, const RangedLP& S, const std::vector<double>& Dr,
          const std::vector<double>& Dc, int max_it, double tol, int gpu_delay_ms) {
    Race out;
    std::atomic<bool> stop{false};
    std::atomic<int> win{-1};
    auto claim = [&](int who) {
        int none = -1;
        if (win.compare_exchange_strong(none, who)) stop.store(true, std::memory_order_relaxed);
    };
    std::thread cpu([&] {
        const auto t0 = Clock::now();
        try {
            out.cpu = cpu_solve(U, S, Dr, Dc, max_it, tol, &stop);
            out.cpu_ok = true;
            if (out.cpu.converged) claim(0);
        } catch (const std::exception& e) {
            out.cpu_err = e.what();
        }
        out.cpu_s = since(t0);
    });
    std::thread gpu([&] {
        if (gpu_delay_ms > 0) std::this_thread::sleep_for(std::chrono::milliseconds(gpu_delay_ms));
        const auto t0 = Clock::now();
        try {
            out.gpu = solve_pdhg_gpu(U, S, Dr, Dc, 0.99, max_it, 64, tol, &stop);
            out.gpu_ok = true;
            if (out.gpu.converged) claim(1);
        } catch (const std::exception& e) {
            out.gpu_err = e.what();
        }
        out.gpu_s = since(t0);
    });
    cpu.join();
    gpu.join();
    out.winner = win.load();
    return out;
}

RangedLP transport(int S, int D) {
    RangedLP lp;
    const int n = S * D, m = S + D, nnz = 2 * n;
    lp.A.rows = m; lp.A.cols = n;
    lp.A.row_idx.resize(nnz); lp.A.col_idx.resize(nnz); lp.A.val.resize(nnz);
    lp.c.resize(n); lp.l.resize(n); lp.u.resize(n); lp.rL.resize(m); lp.rU.resize(m);
    suprema::check(supr_gen_transport(S, D, 7, 100, lp.A.row_idx.data(), lp.A.col_idx.data(),
                                      lp.A.val.data(), lp.c.data(), lp.rL.data(), lp.rU.data(),
                                      lp.l.data(), lp.u.data()), "gen");
    lp.col_names.resize(n); lp.row_names.resize(m);
    return lp;
}

double worst(const KKTReport& k) { return std::fmax(k.eps_P, std::fmax(k.eps_D, k.eps_G)); }

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: race <libsovref.so>\n"); return 2; }
    void* h = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    if (!h || !(g_ref = reinterpret_cast<RefPdlp>(dlsym(h, "ref_pdlp_live")))) {
        std::fprintf(stderr, "cannot load ref_pdlp_live from %s\n", argv[1]);
        return 2;
    }
    const double tol = 1e-6;
    int failures = 0;

    // 1. the GPU wins on a 300,000-variable LP
    RangedLP big = transport(1000, 300);
    auto fb = build_pdlp_form(big);
    ScalingResult sb = scale(fb.first);
    Race r1 = race(fb.first, sb.scaled, sb.Dr, sb.Dc, 500000, tol, 0);
    double eps1 = INFINITY;
    if (r1.winner == 1) {
        std::vector<double> x, y;
        extract_solution(fb.second, r1.gpu.x, r1.gpu.y, x, y);
        eps1 = worst(verify(big, x, y));
    }
    failures += !(r1.winner == 1 && r1.gpu_ok && r1.gpu.converged && eps1 <= tol);
    failures += !(r1.cpu_ok && !r1.cpu.converged && r1.cpu_s < r1.gpu_s + 1.0);

    // 2. a staggered GPU start on a small LP: the CPU wins, the GPU stops at once
    RangedLP small = transport(20, 15);
    auto fs = build_pdlp_form(small);
    ScalingResult ss = scale(fs.first);
    Race r2 = race(fs.first, ss.scaled, ss.Dr, ss.Dc, 500000, tol, 300);
    double eps2 = INFINITY;
    if (r2.winner == 0) {
        std::vector<double> x, y;
        extract_solution(fs.second, r2.cpu.x, r2.cpu.y, x, y);
        eps2 = worst(verify(small, x, y));
    }
    failures += !(r2.winner == 0 && r2.cpu.converged && eps2 <= tol);
    failures += !(r2.gpu_ok && !r2.gpu.converged && r2.gpu_s < 0.5);

    std::printf("{\"gpu_wins\": {\"winner\": %d, \"gpu_s\": %.4f, \"gpu_attempts\": %d, \"verified_eps\": %.3e, "
                "\"cpu_s\": %.4f, \"cpu_attempts\": %d, \"cpu_converged\": %d}, "
                "\"cpu_wins\": {\"winner\": %d, \"cpu_s\": %.4f, \"cpu_attempts\": %d, \"verified_eps\": %.3e, "
                "\"gpu_s\": %.4f, \"gpu_attempts\": %d, \"gpu_converged\": %d}, \"errors\": [\"%s\", \"%s\", \"%s\", \"%s\"], "
                "\"failures\": %d}\n",
                r1.winner, r1.gpu_s, r1.gpu.iterations, eps1, r1.cpu_s, r1.cpu.iterations, (int)r1.cpu.converged,
                r2.winner, r2.cpu_s, r2.cpu.iterations, eps2, r2.gpu_s, r2.gpu.iterations, (int)r2.gpu.converged,
                r1.cpu_err.c_str(), r1.gpu_err.c_str(), r2.cpu_err.c_str(), r2.gpu_err.c_str(), failures);
    return failures == 0 ? 0 : 1;
}
