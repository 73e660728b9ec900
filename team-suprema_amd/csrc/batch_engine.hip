// batch_engine.hip -- many independent LPs solved together (supr_batch_*).
//
// Host side: concatenate the instances block-diagonally once; per solve,
// lay the requested instances out over the launches (block -> instance
// tables), cold-start them, and replay a captured graph of G attempts plus
// the KKT-check group until every instance has converged or hit its limit.
// When the number of running instances has halved, the finished ones are
// dropped from the layout and the graph is captured again (compaction); an
// instance keeps its block counts, so its arithmetic does not change.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "batch_kernels.hpp"
#include "host_lp.hpp"
#include "suprema_pdlp.h"

namespace supr {
namespace {

struct HipError : std::runtime_error {
    explicit HipError(const std::string& m) : std::runtime_error(m) {}
};
struct ArgError : std::runtime_error {
    explicit ArgError(const std::string& m) : std::runtime_error(m) {}
};

#define BATCH_HIP(expr)                                                                    \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            throw HipError(std::string("HIP error at " __FILE__ ":") +                    \
                           std::to_string(__LINE__) + " (" #expr ") -- " +               \
                           hipGetErrorString(_e));                                         \
    } while (0)

class Buf {
   public:
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { if (p_) (void)hipFree(p_); }
    template <typename T> T* alloc(size_t count) {
        if (p_) { (void)hipFree(p_); p_ = nullptr; }
        BATCH_HIP(hipMalloc(&p_, std::max<size_t>(count * sizeof(T), 256)));
        return static_cast<T*>(p_);
    }
    template <typename T> T* put(const std::vector<T>& h, hipStream_t s) {
        T* d = alloc<T>(h.size());
        if (!h.empty()) {
            BATCH_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
            BATCH_HIP(hipStreamSynchronize(s));
        }
        return d;
    }
    template <typename T> T* get() const { return static_cast<T*>(p_); }

   private:
    void* p_ = nullptr;
};

// Lanes per row from the mean row length (times the configured bias).
int lanes_for(double mean_len, double bias) {
    const double t = mean_len * bias;
    if (t < 3.0) return 1;
    if (t < 12.0) return 4;
    if (t < 48.0) return 16;
    return 64;
}

// Blocks of one side: one pass over `rows` rows at L lanes per row, at
// least one block and at most `cap`.
int blocks_for(int rows, int L, int cap) {
    const long need = ((long)rows * L + kBlock - 1) / kBlock;
    return (int)std::max(1L, std::min<long>(need, cap));
}

double inf_norm_finite(const double* v, int len) {
    double m = 0.0;
    for (int k = 0; k < len; ++k)
        if (std::isfinite(v[k])) m = std::max(m, std::fabs(v[k]));
    return m;
}

}  // namespace

struct Batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int count = 0;
    bool mixed = true;
    supr_pdlp_options opt{};
    std::vector<BatchInst> inst;
    std::vector<double> meanA, meanAT;
    std::vector<DevState> init;   // per-instance initial state (max_iter set per solve)
    std::vector<double> c_unscaled, obj_offset;   // for the reported objective
    int64_t M = 0, N = 0, Z = 0;
    double bias = 1.0;
    int cap = 32;
    int64_t BNall = 0, BMall = 0;  // block totals of the whole batch at this configuration
    // device storage
    Buf ptrA, idxA, ptrAT, idxAT, uA, uAT, dA, dAT, fA, fAT;
    Buf fx[2], fy[2], fax[2], faty, fxsum, fysum, fxlast, fylast, fxd, fc, fl, fu, fb;
    Buf dx[2], dy[2], dax[2], daty, dxsum, dysum, dxlast, dylast, dc, dl, du, db;
    Buf ineq, cu, lu, uu, rLu, rUu, Dr, Dc, xu, xu2, xavg, yu, yu2, yavg, atyavg, axavg, bestx, besty;
    Buf dstate, ddx2, dparts, dstatus, dinst, dtabn, dtabm, dids;
    BatchVals V{};
    IterBufs<float> F{};
    IterBufs<double> D{};
    KktBufs K{};
    double *part_n = nullptr, *part_m = nullptr, *part_kn = nullptr, *part_km = nullptr;
    double *part_rn = nullptr, *part_rm = nullptr;

    ~Batch() { if (stream) (void)hipStreamDestroy(stream); }

    void build(int32_t cnt, const supr_lp* U, const supr_lp* S, const double* const* Drs,
               const double* const* Dcs, const supr_pdlp_options& o) {
        count = cnt;
        opt = o;
        mixed = o.mode != SUPR_MODE_FAST_FP64;
        BATCH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        inst.assign(cnt, BatchInst{});
        meanA.resize(cnt);
        meanAT.resize(cnt);
        init.resize(cnt);
        obj_offset.resize(cnt);
        for (int k = 0; k < cnt; ++k) { M += S[k].m; N += S[k].n; Z += S[k].nnz; }
        std::vector<int32_t> hptrA, hptrAT, hidxA(Z), hidxAT(Z);
        std::vector<double> hvA(Z), hvAT(Z), hvuA(Z), hvuAT(Z);
        hptrA.reserve(M + cnt);
        hptrAT.reserve(N + cnt);
        std::vector<double> c(N), l(N), u(N), b(M), hlu(N), huu(N), hrL(M), hrU(M), hDr(M), hDc(N);
        c_unscaled.assign(N, 0.0);
        std::vector<uint8_t> hineq(M);
        int64_t ro = 0, co = 0, za = 0;
        for (int k = 0; k < cnt; ++k) {
            const supr_lp& s = S[k];
            const supr_lp& un = U[k];
            BatchInst& I = inst[k];
            I.m = s.m;
            I.n = s.n;
            I.ro = ro;
            I.co = co;
            I.pa = (int64_t)hptrA.size();
            I.pat = (int64_t)hptrAT.size();
            I.za = za;
            const Csr A = build_csr(s, s.val, false), AT = build_csr(s, s.val, true);
            hptrA.insert(hptrA.end(), A.ptr.begin(), A.ptr.end());
            hptrAT.insert(hptrAT.end(), AT.ptr.begin(), AT.ptr.end());
            for (int32_t p = 0; p < s.nnz; ++p) {
                hidxA[za + p] = A.idx[p];
                hvA[za + p] = A.val[p];
                hvuA[za + p] = un.val[A.perm[p]];
                hidxAT[za + p] = AT.idx[p];
                hvAT[za + p] = AT.val[p];
                hvuAT[za + p] = un.val[AT.perm[p]];
            }
            meanA[k] = (double)s.nnz / s.m;
            meanAT[k] = (double)s.nnz / s.n;
            for (int j = 0; j < s.n; ++j) {
                c[co + j] = s.c[j]; l[co + j] = s.l[j]; u[co + j] = s.u[j];
                c_unscaled[co + j] = un.c[j]; hlu[co + j] = un.l[j]; huu[co + j] = un.u[j];
                hDc[co + j] = Dcs[k][j];
            }
            for (int i = 0; i < s.m; ++i) {
                b[ro + i] = s.rL[i];
                hineq[ro + i] = s.rL[i] != s.rU[i];
                hrL[ro + i] = un.rL[i]; hrU[ro + i] = un.rU[i];
                hDr[ro + i] = Drs[k][i];
            }
            init[k] = initial_state(s, un);
            obj_offset[k] = un.obj_offset;
            ro += s.m;
            co += s.n;
            za += s.nnz;
        }
        hipStream_t st = stream;
        ptrA.put(hptrA, st);
        idxA.put(hidxA, st);
        ptrAT.put(hptrAT, st);
        idxAT.put(hidxAT, st);
        V.uA = uA.put(hvuA, st);
        V.uAT = uAT.put(hvuAT, st);
        V.dA = dA.put(hvA, st);
        V.dAT = dAT.put(hvAT, st);
        auto iter = [&](auto& B, auto tag, Buf* x, Buf* y, Buf* ax, Buf& aty, Buf& xsum, Buf& ysum,
                        Buf& xlast, Buf& ylast, Buf& cb, Buf& lb, Buf& ub, Buf& bb) {
            using T = decltype(tag);
            for (int q = 0; q < 2; ++q) {
                B.x[q] = x[q].alloc<T>(N);
                B.y[q] = y[q].alloc<T>(M);
                B.ax[q] = ax[q].alloc<T>(M);
            }
            B.aty = aty.alloc<T>(N);
            B.xsum = xsum.alloc<T>(N);
            B.ysum = ysum.alloc<T>(M);
            B.xlast = xlast.alloc<T>(N);
            B.ylast = ylast.alloc<T>(M);
            B.c = cb.put(std::vector<T>(c.begin(), c.end()), st);
            B.l = lb.put(std::vector<T>(l.begin(), l.end()), st);
            B.u = ub.put(std::vector<T>(u.begin(), u.end()), st);
            B.b = bb.put(std::vector<T>(b.begin(), b.end()), st);
        };
        const uint8_t* dineq = ineq.put(hineq, st);
        iter(D, double{}, dx, dy, dax, daty, dxsum, dysum, dxlast, dylast, dc, dl, du, db);
        D.xd = nullptr;
        D.ineq = dineq;
        if (mixed) {
            V.fA = fA.put(std::vector<float>(hvA.begin(), hvA.end()), st);
            V.fAT = fAT.put(std::vector<float>(hvAT.begin(), hvAT.end()), st);
            iter(F, float{}, fx, fy, fax, faty, fxsum, fysum, fxlast, fylast, fc, fl, fu, fb);
            F.xd = fxd.alloc<float>(2 * N);
            F.ineq = dineq;
        } else {
            F = IterBufs<float>{};
            F.xd = nullptr;
        }
        K.c = cu.put(c_unscaled, st); K.l = lu.put(hlu, st); K.u = uu.put(huu, st);
        K.rL = rLu.put(hrL, st); K.rU = rUu.put(hrU, st);
        K.Dr = Dr.put(hDr, st); K.Dc = Dc.put(hDc, st);
        K.xu = xu.alloc<double>(N); K.xu2 = xu2.alloc<double>(N); K.xavg = xavg.alloc<double>(N);
        K.aty_avg = atyavg.alloc<double>(N); K.best_x = bestx.alloc<double>(N);
        K.yu = yu.alloc<double>(M); K.yu2 = yu2.alloc<double>(M); K.yavg = yavg.alloc<double>(M);
        K.ax_avg = axavg.alloc<double>(M); K.best_y = besty.alloc<double>(M);
        dstate.alloc<DevState>(2 * (size_t)cnt);
        ddx2.alloc<double>(cnt);
        dstatus.alloc<double>(2);
        dinst.alloc<BatchInst>(cnt);
        dids.alloc<int32_t>(cnt);
        set_config(1.0, 32);
    }

    // include/pdlp_algo.hpp:67-77 starting values and the single-LP
    // engine's control fields (Solver::initial_state).
    DevState initial_state(const supr_lp& s, const supr_lp& un) const {
        DevState st;
        std::memset(&st, 0, sizeof(st));
        double amax = 0.0, cn = 0.0, bn = 0.0;
        for (int32_t p = 0; p < s.nnz; ++p) amax = smax(amax, std::fabs(s.val[p]));
        for (int j = 0; j < s.n; ++j) cn += s.c[j] * s.c[j];
        for (int i = 0; i < s.m; ++i) if (std::isfinite(s.rL[i])) bn += s.rL[i] * s.rL[i];
        cn = std::sqrt(cn);
        bn = std::sqrt(bn);
        st.eta = amax > 0 ? 1.0 / amax : 1.0;
        st.omega = (cn > 1e-10 && bn > 1e-10) ? cn / bn : 1.0;
        st.merit_last_restart = INFINITY;
        st.merit_prev_cand = INFINITY;
        st.best_merit = INFINITY;
        st.stall_ref = INFINITY;
        st.tol = opt.tol;
        double sw = opt.switch_tol;
        if (sw <= 0) sw = opt.tol < 1e-5 ? 1e-5 : 0.0;
        st.switch_tol = sw;
        st.stall_checks = opt.stall_checks > 0 ? opt.stall_checks : 6;
        st.denom_p = 1.0 + std::max(inf_norm_finite(un.rL, un.m), inf_norm_finite(un.rU, un.m));
        st.denom_d = 1.0 + inf_norm_finite(un.c, un.n);
        st.obj_offset = un.obj_offset;
        st.check_every = opt.check_every;
        st.need_aty = 1;
        st.phase = mixed ? 0 : 1;
        return st;
    }

    // Lanes rule and per-instance block cap; partial-sum storage for the
    // whole batch at this configuration.
    void set_config(double b, int c) {
        int64_t bn_all = 0, bm_all = 0;
        std::vector<BatchInst> next(inst);
        for (int k = 0; k < count; ++k) {
            BatchInst& I = next[k];
            I.LA = lanes_for(meanA[k], b);
            I.LAT = lanes_for(meanAT[k], b);
            I.bn = blocks_for(I.n, I.LAT, c);
            I.bm = blocks_for(I.m, I.LA, c);
            bn_all += I.bn;
            bm_all += I.bm;
        }
        if (bn_all + bm_all > (int64_t)INT32_MAX / 16) throw ArgError("batch too large for one launch");
        double* p = dparts.alloc<double>((size_t)10 * bn_all + 9 * bm_all);
        part_n = p; p += bn_all;
        part_kn = p; p += 8 * bn_all;
        part_rn = p; p += bn_all;
        part_m = p; p += 2 * bm_all;
        part_km = p; p += 6 * bm_all;
        part_rm = p;
        dtabn.alloc<int32_t>(bn_all);
        dtabm.alloc<int32_t>(bm_all);
        inst.swap(next);
        BNall = bn_all;
        BMall = bm_all;
        bias = b;
        cap = c;
    }

    // Lays the instances `live` out over the launches; returns the geometry.
    BatchGeo layout(const std::vector<int32_t>& live) {
        std::vector<int32_t> tn, tm;
        int32_t on = 0, om = 0;
        for (int32_t k : live) {
            BatchInst& I = inst[k];
            I.on = on;
            I.om = om;
            tn.insert(tn.end(), I.bn, k);
            tm.insert(tm.end(), I.bm, k);
            on += I.bn;
            om += I.bm;
        }
        hipStream_t st = stream;
        BATCH_HIP(hipMemcpyAsync(dinst.get<BatchInst>(), inst.data(), sizeof(BatchInst) * count,
                                 hipMemcpyHostToDevice, st));
        BATCH_HIP(hipMemcpyAsync(dtabn.get<int32_t>(), tn.data(), sizeof(int32_t) * tn.size(),
                                 hipMemcpyHostToDevice, st));
        BATCH_HIP(hipMemcpyAsync(dtabm.get<int32_t>(), tm.data(), sizeof(int32_t) * tm.size(),
                                 hipMemcpyHostToDevice, st));
        BATCH_HIP(hipMemcpyAsync(dids.get<int32_t>(), live.data(), sizeof(int32_t) * live.size(),
                                 hipMemcpyHostToDevice, st));
        BATCH_HIP(hipStreamSynchronize(st));   // the host vectors go out of scope
        BatchGeo g;
        g.inst = dinst.get<BatchInst>();
        g.tab_n = dtabn.get<int32_t>();
        g.tab_m = dtabm.get<int32_t>();
        g.ids = dids.get<int32_t>();
        g.BN = on;
        g.BM = om;
        g.ptrA = ptrA.get<int32_t>();
        g.idxA = idxA.get<int32_t>();
        g.ptrAT = ptrAT.get<int32_t>();
        g.idxAT = idxAT.get<int32_t>();
        return g;
    }

    void enqueue_replay(const BatchGeo& g, int live, int G) {
        hipStream_t st = stream;
        DevState* S = dstate.get<DevState>();
        double* dx2 = ddx2.get<double>();
        for (int a = 0; a < G; ++a) {
            const int p = a & 1;
            hipLaunchKernelGGL(kb_primal, dim3(g.BN), dim3(kBlock), 0, st, S, p, g, V, F, D, part_m, dx2, part_n);
            hipLaunchKernelGGL(kb_dual, dim3(g.BM), dim3(kBlock), 0, st, S, p ^ 1, g, V, F, D, part_n, dx2, part_m);
        }
        hipLaunchKernelGGL(kb_check_prep, dim3(g.BN + g.BM), dim3(kBlock), 0, st, S, g, F, D, K, part_m, dx2);
        hipLaunchKernelGGL(kb_check_kkt, dim3(g.BN + g.BM), dim3(kBlock), 0, st, S, g, V, K, part_kn, part_km);
        hipLaunchKernelGGL(kb_kkt_fin, dim3(live), dim3(kBlock), 0, st, S, g, part_kn, part_km);
        hipLaunchKernelGGL(kb_check_restart, dim3(g.BN + g.BM), dim3(kBlock), 0, st, S, g, F, D, K, part_rn, part_rm);
        if (mixed)
            hipLaunchKernelGGL(kb_switch_products, dim3(g.BN + g.BM), dim3(kBlock), 0, st, S, g, V, D);
        hipLaunchKernelGGL(kb_restart_fin, dim3(live), dim3(kBlock), 0, st, S, g, part_rn, part_rm);
        hipLaunchKernelGGL(kb_status, dim3(1), dim3(kBlock), 0, st, S, g, live, dstatus.get<double>());
        BATCH_HIP(hipGetLastError());
    }

    hipGraphExec_t capture(const BatchGeo& g, int live, int G) {
        hipGraph_t gr = nullptr;
        hipGraphExec_t exec = nullptr;
        BATCH_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_replay(g, live, G);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
        }
        BATCH_HIP(hipStreamEndCapture(stream, &gr));
        const hipError_t e = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        BATCH_HIP(e);
        return exec;
    }

    // Solves the instances `ids` from the cold start; fills states[k] (S[2k]
    // at the end) for k in ids.  Returns seconds.
    double run(const std::vector<int32_t>& ids, int max_iter, std::vector<DevState>& states) {
        hipStream_t st = stream;
        std::vector<DevState> h(2 * (size_t)count);
        {
            std::vector<char> seen(count, 0);
            for (int32_t k : ids) {
                if (k < 0 || k >= count) throw ArgError("instance id out of range");
                if (seen[k]) throw ArgError("instance id listed twice");
                seen[k] = 1;
                DevState s0 = init[k];
                s0.max_iter = max_iter;
                s0.active = max_iter > 0 ? 1 : 0;
                s0.need_check = max_iter > 0 ? 0 : 1;
                h[2 * k] = h[2 * k + 1] = s0;
            }
        }
        const auto t0 = std::chrono::steady_clock::now();
        BATCH_HIP(hipMemcpyAsync(dstate.get<DevState>(), h.data(), sizeof(DevState) * h.size(),
                                 hipMemcpyHostToDevice, st));
        std::vector<int32_t> live(ids);
        BatchGeo g = layout(live);
        if (mixed) {
            hipLaunchKernelGGL(kb_cold<float>, dim3(g.BN + g.BM), dim3(kBlock), 0, st, g, F);
        } else {
            hipLaunchKernelGGL(kb_cold<double>, dim3(g.BN + g.BM), dim3(kBlock), 0, st, g, D);
            hipLaunchKernelGGL(kb_cold_ax<double>, dim3(g.BM), dim3(kBlock), 0, st, g, V, D);
        }
        BATCH_HIP(hipGetLastError());
        int G = opt.graph_attempts > 0 ? opt.graph_attempts : 16;
        G += G & 1;   // the state double buffer needs an even number of attempts per replay
        hipGraphExec_t exec = nullptr;
        constexpr int R = 4;
        double* ring = nullptr;
        hipEvent_t ev[R] = {};
        auto cleanup = [&]() {
            (void)hipStreamSynchronize(st);
            for (auto& e : ev) if (e) { (void)hipEventDestroy(e); e = nullptr; }
            if (ring) { (void)hipHostFree(ring); ring = nullptr; }
            if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
        };
        try {
            BATCH_HIP(hipHostMalloc((void**)&ring, sizeof(double) * 2 * R, hipHostMallocDefault));
            for (auto& e : ev) BATCH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            exec = capture(g, (int)live.size(), G);
            double last_progress = -1.0;
            int idle = 0, replays = 0, inflight = 0;
            size_t live_at_capture = live.size();
            while (true) {
                if (inflight == R) {
                    const int slot = (replays - inflight) % R;
                    BATCH_HIP(hipEventSynchronize(ev[slot]));
                    --inflight;
                    const double running = ring[2 * slot], progress = ring[2 * slot + 1];
                    if (running == 0.0) break;
                    // A replay always advances attempts or checks of a running
                    // instance; a loop that stops moving is a logic error.
                    if (progress == last_progress) {
                        if (++idle > 64) throw std::runtime_error("batched device loop made no progress");
                    } else {
                        idle = 0;
                    }
                    last_progress = progress;
                    if (running * 2 <= (double)live_at_capture && live_at_capture > 8) {
                        // compaction: drain, drop finished instances, capture again
                        BATCH_HIP(hipStreamSynchronize(st));
                        inflight = 0;
                        BATCH_HIP(hipMemcpy(h.data(), dstate.get<DevState>(), sizeof(DevState) * h.size(),
                                            hipMemcpyDeviceToHost));
                        std::vector<int32_t> still;
                        for (int32_t k : live) if (!h[2 * k].done) still.push_back(k);
                        if (still.empty()) break;
                        if (opt.verbose)
                            std::fprintf(stderr, "[suprema] batch: replay %d, %zu of %zu instances running, layout compacted\n",
                                         replays, still.size(), live.size());
                        live.swap(still);
                        live_at_capture = live.size();
                        (void)hipGraphExecDestroy(exec);
                        exec = nullptr;
                        g = layout(live);
                        exec = capture(g, (int)live.size(), G);
                    }
                }
                const int slot = replays % R;
                BATCH_HIP(hipGraphLaunch(exec, st));
                BATCH_HIP(hipMemcpyAsync(&ring[2 * slot], dstatus.get<void>(), sizeof(double) * 2,
                                         hipMemcpyDeviceToHost, st));
                BATCH_HIP(hipEventRecord(ev[slot], st));
                ++replays;
                ++inflight;
            }
            BATCH_HIP(hipStreamSynchronize(st));
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        BATCH_HIP(hipMemcpy(h.data(), dstate.get<DevState>(), sizeof(DevState) * h.size(), hipMemcpyDeviceToHost));
        states.resize(count);
        for (int32_t k : ids) states[k] = h[2 * k];
        return secs;
    }

    // Times every (lanes bias, block cap) candidate on a short run of the
    // whole batch and keeps the fastest.
    void tune(int probe_iters) {
        std::vector<int32_t> all(count);
        std::iota(all.begin(), all.end(), 0);
        std::vector<DevState> scratch_states;
        double best = std::numeric_limits<double>::infinity(), best_bias = 1.0;
        int best_cap = 32;
        for (double b : {0.5, 1.0, 2.0}) {
            for (int c : {8, 32, 128}) {
                set_config(b, c);
                run(all, probe_iters, scratch_states);            // warm
                const double t = run(all, probe_iters, scratch_states);
                if (t < best) { best = t; best_bias = b; best_cap = c; }
            }
        }
        set_config(best_bias, best_cap);
    }
};

}  // namespace supr

struct supr_batch {
    supr::Batch b;
};

extern "C" {

int supr_batch_create(int32_t count, const supr_lp* unscaled_forms, const supr_lp* scaled_forms,
                      const double* const* Dr, const double* const* Dc,
                      const supr_pdlp_options* opt, supr_batch** out) {
    using namespace supr;
    if (!out || count < 1 || !unscaled_forms || !scaled_forms || !Dr || !Dc || !opt) {
        set_error("supr_batch_create: null argument or count < 1");
        return SUPR_EINVAL;
    }
    *out = nullptr;
    if (opt->mode == SUPR_MODE_PLAIN_FP64) {
        set_error("supr_batch_create: the batched engine runs the device-resident modes only");
        return SUPR_EINVAL;
    }
    if (opt->max_iterations < 0 || opt->check_every < 1 || !(opt->tol >= 0)) {
        set_error("supr_batch_create: bad options");
        return SUPR_EINVAL;
    }
    int64_t zsum = 0;
    for (int k = 0; k < count; ++k) {
        const supr_lp &U = unscaled_forms[k], &S = scaled_forms[k];
        if (!check_lp(&U, "unscaled_form") || !check_lp(&S, "scaled_form")) return SUPR_EINVAL;
        if (U.m != S.m || U.n != S.n || U.nnz != S.nnz || S.m < 1 || S.n < 1 || !Dr[k] || !Dc[k]) {
            set_error("supr_batch_create: instance " + std::to_string(k) + " has mismatched or empty forms");
            return SUPR_EINVAL;
        }
        for (int32_t p = 0; p < S.nnz; ++p)
            if (U.row_idx[p] != S.row_idx[p] || U.col_idx[p] != S.col_idx[p]) {
                set_error("supr_batch_create: instance " + std::to_string(k) + " forms differ in pattern");
                return SUPR_EINVAL;
            }
        zsum += S.nnz;
    }
    if (zsum > INT32_MAX) { set_error("supr_batch_create: total nnz exceeds int32"); return SUPR_EINVAL; }
    int devs = 0;
    if (hipGetDeviceCount(&devs) != hipSuccess || devs < 1) { set_error("no HIP device visible"); return SUPR_ENODEV; }
    try {
        BATCH_HIP(hipSetDevice(opt->device));
        auto h = std::make_unique<supr_batch>();
        h->b.device = opt->device;
        h->b.build(count, unscaled_forms, scaled_forms, Dr, Dc, *opt);
        if (opt->spmv_lanes_A == 0) h->b.tune(128);
        *out = h.release();
        return SUPR_OK;
    } catch (const ArgError& e) {
        set_error(e.what());
        return SUPR_EINVAL;
    } catch (const HipError& e) {
        set_error(e.what());
        return SUPR_EDEVICE;
    } catch (const std::bad_alloc&) {
        set_error("out of host memory");
        return SUPR_ENOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SUPR_EINTERNAL;
    }
}

int supr_batch_configure(supr_batch* batch, double lanes_bias, int32_t max_blocks) {
    using namespace supr;
    if (!batch || !(lanes_bias > 0) || max_blocks < 1 || max_blocks > kMaxBlocks) {
        set_error("supr_batch_configure: need lanes_bias > 0 and 1 <= max_blocks <= 2048");
        return SUPR_EINVAL;
    }
    try {
        BATCH_HIP(hipSetDevice(batch->b.device));
        batch->b.set_config(lanes_bias, max_blocks);
        return SUPR_OK;
    } catch (const ArgError& e) {
        set_error(e.what());
        return SUPR_EINVAL;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SUPR_EDEVICE;
    }
}

int supr_batch_config(const supr_batch* batch, double* lanes_bias, int32_t* max_blocks) {
    if (!batch || !lanes_bias || !max_blocks) { supr::set_error("null argument"); return SUPR_EINVAL; }
    *lanes_bias = batch->b.bias;
    *max_blocks = batch->b.cap;
    return SUPR_OK;
}

int supr_batch_solve(supr_batch* batch, const int32_t* ids, int32_t nids, double* const* x_out,
                     double* const* y_out, supr_pdlp_info* infos, double* seconds) {
    using namespace supr;
    if (!batch || (ids && nids < 1) || !infos) {
        set_error("supr_batch_solve: null argument or empty id list");
        return SUPR_EINVAL;
    }
    try {
        Batch& b = batch->b;
        BATCH_HIP(hipSetDevice(b.device));
        std::vector<int32_t> list;
        if (ids) list.assign(ids, ids + nids);
        else { list.resize(b.count); std::iota(list.begin(), list.end(), 0); }
        std::vector<DevState> st;
        const double secs = b.run(list, b.opt.max_iterations, st);
        if (seconds) *seconds = secs;
        std::vector<double> xh;
        for (size_t r = 0; r < list.size(); ++r) {
            const int32_t id = list[r];
            const BatchInst& I = b.inst[id];
            const DevState& s = st[id];
            supr_pdlp_info& inf = infos[r];
            std::memset(&inf, 0, sizeof(inf));
            inf.iterations = s.attempts;
            inf.restarts = s.restarts;
            inf.converged = s.done == 1 ? 1 : 0;
            inf.has_point = s.has_best ? 1 : 0;
            inf.eps_P = s.best_eps[0]; inf.eps_D = s.best_eps[1]; inf.eps_G = s.best_eps[2];
            inf.iterations_fp32 = b.mixed ? (s.phase == 0 ? s.attempts : s.attempts_fp32) : 0;
            inf.checks = s.checks;
            inf.lanes_A = I.LA; inf.lanes_AT = I.LAT;
            inf.seconds_solve = secs;
            double obj = b.obj_offset[id];
            if (s.has_best) {
                // the returned point and its objective c'x + offset (unscaled form)
                xh.resize(I.n);
                BATCH_HIP(hipMemcpy(xh.data(), b.K.best_x + I.co, sizeof(double) * I.n, hipMemcpyDeviceToHost));
                double cx = 0.0;
                for (int j = 0; j < I.n; ++j) cx += b.c_unscaled[I.co + j] * xh[j];
                obj = cx + b.obj_offset[id];
                if (x_out && x_out[r]) std::copy(xh.begin(), xh.end(), x_out[r]);
                if (y_out && y_out[r])
                    BATCH_HIP(hipMemcpy(y_out[r], b.K.best_y + I.ro, sizeof(double) * I.m, hipMemcpyDeviceToHost));
            }
            inf.objective = obj;
        }
        return SUPR_OK;
    } catch (const ArgError& e) {
        set_error(e.what());
        return SUPR_EINVAL;
    } catch (const HipError& e) {
        set_error(e.what());
        return SUPR_EDEVICE;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SUPR_EINTERNAL;
    }
}

void supr_batch_destroy(supr_batch* batch) {
    if (!batch) return;
    (void)hipSetDevice(batch->b.device);
    delete batch;
}

}  // extern "C"
