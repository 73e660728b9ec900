// batch_engine.hip -- many independent LPs solved together: the batched
// engine (supr_batch_*) and the serving engine (supr_server_*).
//
// Both sit on one resident pool.  The pool holds fixed-capacity device
// arenas -- one index space for columns (every per-column vector lives at
// the same offset in each column array), one for rows, one for nonzeros --
// and a fixed number of state slots.  An instance occupies a column range
// (n + 1 entries: its A^T row pointers share the offset), a row range
// (m + 1: its A row pointers), a nonzero range and a slot, taken first-fit;
// admission is whether those ranges fit.
//
// Running instances are laid out over the launches of the segmented
// device-resident loop (batch_kernels.hpp) and a graph of G attempts plus
// the KKT-check group is replayed.  Membership changes only at replay
// boundaries: the batch drops finished instances when half have finished
// (compaction); the server retires finished instances and admits queued
// ones after every step.  An instance's block counts and lanes are fixed at
// admission, so its arithmetic never depends on its neighbours.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "batch_kernels.hpp"
#include "host_lp.hpp"
#include "suprema_pdlp.h"

namespace supr {
namespace {

using Clock = std::chrono::steady_clock;

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
    template <typename T> T* get() const { return static_cast<T*>(p_); }

   private:
    void* p_ = nullptr;
};

// First-fit interval allocator over [0, cap) with coalescing frees.
class Arena {
   public:
    void reset(int64_t cap) {
        free_.clear();
        if (cap > 0) free_[0] = cap;
    }
    bool alloc(int64_t len, int64_t& off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < len) continue;
            off = it->first;
            const int64_t rest = it->second - len;
            free_.erase(it);
            if (rest) free_[off + len] = rest;
            return true;
        }
        return false;
    }
    void release(int64_t off, int64_t len) {
        auto it = free_.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

   private:
    std::map<int64_t, int64_t> free_;
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

// A side of an instance runs the LDS-staged tiles (256 row products staged,
// coalesced epilogues) when it has L >= 2 and at least this many rows.
constexpr int kBatchStageRows = 1024;
int staged_side(int rows, int L) { return (L >= 2 && rows >= kBatchStageRows) ? 1 : 0; }

// Blocks of a staged side: one 256-row tile each, at most `cap`.
int tiles_for(int rows, int cap) {
    const long need = ((long)rows + kBlock - 1) / kBlock;
    return (int)std::max(1L, std::min<long>(need, cap));
}

double inf_norm_finite(const double* v, int len) {
    double m = 0.0;
    for (int k = 0; k < len; ++k)
        if (std::isfinite(v[k])) m = std::max(m, std::fabs(v[k]));
    return m;
}

// Device bytes per column, row and nonzero of a resident instance: the
// FP64 iterate arrays, the KKT-check arrays, the row pointers; in the
// mixed mode also the FP32 iterate arrays and FP32 values.
constexpr int64_t kColBytes64 = 8 * 8 + 9 * 8 + 4, kColBytes32 = 8 * 4 + 8;
constexpr int64_t kRowBytes64 = 7 * 8 + 1 + 8 * 8 + 4, kRowBytes32 = 7 * 4;
constexpr int64_t kNzBytes64 = 2 * 4 + 4 * 8, kNzBytes32 = 2 * 4;

int64_t footprint_bytes(int64_t m, int64_t n, int64_t nnz, bool mixed) {
    return (n + 1) * (kColBytes64 + (mixed ? kColBytes32 : 0)) +
           (m + 1) * (kRowBytes64 + (mixed ? kRowBytes32 : 0)) +
           nnz * (kNzBytes64 + (mixed ? kNzBytes32 : 0));
}

}  // namespace

// Host-side arrays of one instance, built on the caller's thread.
struct Prepared {
    int32_t m = 0, n = 0, nnz = 0;
    std::vector<int32_t> ptrA, idxA, ptrAT, idxAT;
    std::vector<double> vA, vAT, vuA, vuAT;
    std::vector<float> fA, fAT;
    std::vector<double> c, l, u, b, cu, lu, uu, rLu, rUu, Dr, Dc;
    std::vector<float> fc, fl, fu, fb;
    std::vector<uint8_t> ineq;
    double meanA = 0.0, meanAT = 0.0;
    DevState init{};
};

// include/pdlp_algo.hpp:67-77 starting values and the single-LP engine's
// control fields (Solver::initial_state).
DevState initial_state(const supr_lp& s, const supr_lp& un, const supr_pdlp_options& o, bool mixed) {
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
    st.tol = o.tol;
    double sw = o.switch_tol;
    if (sw <= 0) sw = o.tol < 1e-5 ? 1e-5 : 0.0;
    st.switch_tol = sw;
    st.stall_checks = o.stall_checks > 0 ? o.stall_checks : 6;
    st.denom_p = 1.0 + std::max(inf_norm_finite(un.rL, un.m), inf_norm_finite(un.rU, un.m));
    st.denom_d = 1.0 + inf_norm_finite(un.c, un.n);
    st.obj_offset = un.obj_offset;
    st.check_every = o.check_every;
    st.need_aty = 1;
    st.phase = mixed ? 0 : 1;
    return st;
}

Prepared prepare(const supr_lp& U, const supr_lp& S, const double* Dr, const double* Dc,
                 const supr_pdlp_options& o, bool mixed) {
    Prepared p;
    p.m = S.m;
    p.n = S.n;
    p.nnz = S.nnz;
    Csr A = build_csr(S, S.val, false), AT = build_csr(S, S.val, true);
    p.vuA.resize(S.nnz);
    p.vuAT.resize(S.nnz);
    for (int32_t q = 0; q < S.nnz; ++q) {
        p.vuA[q] = U.val[A.perm[q]];
        p.vuAT[q] = U.val[AT.perm[q]];
    }
    p.ptrA = std::move(A.ptr);
    p.idxA = std::move(A.idx);
    p.vA = std::move(A.val);
    p.ptrAT = std::move(AT.ptr);
    p.idxAT = std::move(AT.idx);
    p.vAT = std::move(AT.val);
    p.c.assign(S.c, S.c + S.n);
    p.l.assign(S.l, S.l + S.n);
    p.u.assign(S.u, S.u + S.n);
    p.b.assign(S.rL, S.rL + S.m);
    p.ineq.resize(S.m);
    for (int i = 0; i < S.m; ++i) p.ineq[i] = S.rL[i] != S.rU[i];
    p.cu.assign(U.c, U.c + U.n);
    p.lu.assign(U.l, U.l + U.n);
    p.uu.assign(U.u, U.u + U.n);
    p.rLu.assign(U.rL, U.rL + U.m);
    p.rUu.assign(U.rU, U.rU + U.m);
    p.Dr.assign(Dr, Dr + S.m);
    p.Dc.assign(Dc, Dc + S.n);
    if (mixed) {
        p.fA.assign(p.vA.begin(), p.vA.end());
        p.fAT.assign(p.vAT.begin(), p.vAT.end());
        p.fc.assign(p.c.begin(), p.c.end());
        p.fl.assign(p.l.begin(), p.l.end());
        p.fu.assign(p.u.begin(), p.u.end());
        p.fb.assign(p.b.begin(), p.b.end());
    }
    p.meanA = S.m > 0 ? (double)S.nnz / S.m : 0.0;
    p.meanAT = (double)S.nnz / S.n;
    p.init = initial_state(S, U, o, mixed);
    return p;
}

// Checks one instance's pair of forms (what supr_batch_create and
// supr_server_submit accept).
bool check_instance(const supr_lp& U, const supr_lp& S, const double* Dr, const double* Dc,
                    const std::string& what) {
    if (!check_lp(&U, "unscaled_form") || !check_lp(&S, "scaled_form")) return false;
    if (U.m != S.m || U.n != S.n || U.nnz != S.nnz || S.m < 0 || S.n < 1 || !Dr || !Dc) {
        set_error(what + " has mismatched or empty forms");
        return false;
    }
    for (int32_t p = 0; p < S.nnz; ++p)
        if (U.row_idx[p] != S.row_idx[p] || U.col_idx[p] != S.col_idx[p]) {
            set_error(what + ": the forms differ in pattern");
            return false;
        }
    return true;
}

struct Pool {
    int device = 0;
    hipStream_t stream = nullptr;
    bool mixed = true;
    supr_pdlp_options opt{};
    int64_t Ncap = 0, Mcap = 0, Zcap = 0;
    int Scap = 0;
    Arena an, am, az;
    std::vector<int32_t> free_slots;
    std::vector<BatchInst> inst;
    std::vector<int64_t> nnz_slot;
    std::vector<DevState> init;
    std::vector<double> meanA, meanAT;
    std::vector<double> obj_offset;
    double bias = 1.0;
    int cap = 32;
    // device storage
    Buf ptrA, idxA, ptrAT, idxAT, uA, uAT, dA, dAT, fA, fAT;
    Buf fx[2], fy[2], fax[2], faty, fxsum, fysum, fxlast, fylast, fxd, fc, fl, fu, fb;
    Buf dx[2], dy[2], dax[2], daty, dxsum, dysum, dxlast, dylast, dc, dl, du, db;
    Buf ineq, cu, lu, uu, rLu, rUu, Dr, Dc, xu, xu2, xavg, yu, yu2, yavg, atyavg, axavg, bestx, besty;
    Buf dstate, ddx2, dparts, dstatus, dinst, dtabn, dtabm, dids, dstage, dslots;
    BatchVals V{};
    IterBufs<float> F{};
    IterBufs<double> D{};
    KktBufs K{};
    double *part_n = nullptr, *part_m = nullptr, *part_kn = nullptr, *part_km = nullptr;
    double *part_rn = nullptr, *part_rm = nullptr;
    int64_t device_bytes = 0;
    double* status_h = nullptr;   // host-coherent [running, progress] written by kb_status
    double* status_d = nullptr;   // the same memory, device address
    int32_t* done_h = nullptr;    // host-coherent done flag per slot, written by kb_restart_fin
    int32_t* done_d = nullptr;

    // uploads of admitted requests (the server): on their own stream, so they
    // overlap the running set's replays on `stream`; null: uploads on `stream`
    hipStream_t upload_stream = nullptr;

    ~Pool() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (status_h) (void)hipHostFree(status_h);
        if (done_h) (void)hipHostFree(done_h);
        if (upload_stream) (void)hipStreamDestroy(upload_stream);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <typename T> T* arr(Buf& b, int64_t len) {
        device_bytes += (int64_t)sizeof(T) * std::max<int64_t>(len, 1);
        return b.alloc<T>((size_t)std::max<int64_t>(len, 1));
    }

    void create(int64_t ncap, int64_t mcap, int64_t zcap, int scap, const supr_pdlp_options& o,
                double b, int c) {
        opt = o;
        mixed = o.mode != SUPR_MODE_FAST_FP64;
        Ncap = ncap; Mcap = mcap; Zcap = zcap; Scap = scap;
        an.reset(Ncap); am.reset(Mcap); az.reset(Zcap);
        free_slots.resize(Scap);
        for (int k = 0; k < Scap; ++k) free_slots[k] = Scap - 1 - k;   // slot 0 is taken first
        inst.assign(Scap, BatchInst{});
        nnz_slot.assign(Scap, 0);
        init.resize(Scap);
        meanA.assign(Scap, 0.0);
        meanAT.assign(Scap, 0.0);
        obj_offset.assign(Scap, 0.0);
        BATCH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        arr<int32_t>(ptrA, Mcap); arr<int32_t>(idxA, Zcap);
        arr<int32_t>(ptrAT, Ncap); arr<int32_t>(idxAT, Zcap);
        V.uA = arr<double>(uA, Zcap); V.uAT = arr<double>(uAT, Zcap);
        V.dA = arr<double>(dA, Zcap); V.dAT = arr<double>(dAT, Zcap);
        auto iter = [&](auto& B, auto tag, Buf* x, Buf* y, Buf* ax, Buf& aty, Buf& xsum, Buf& ysum,
                        Buf& xlast, Buf& ylast, Buf& cb, Buf& lb, Buf& ub, Buf& bb) {
            using T = decltype(tag);
            for (int q = 0; q < 2; ++q) {
                B.x[q] = arr<T>(x[q], Ncap);
                B.y[q] = arr<T>(y[q], Mcap);
                B.ax[q] = arr<T>(ax[q], Mcap);
            }
            B.aty = arr<T>(aty, Ncap);
            B.xsum = arr<T>(xsum, Ncap);
            B.ysum = arr<T>(ysum, Mcap);
            B.xlast = arr<T>(xlast, Ncap);
            B.ylast = arr<T>(ylast, Mcap);
            B.c = arr<T>(cb, Ncap);
            B.l = arr<T>(lb, Ncap);
            B.u = arr<T>(ub, Ncap);
            B.b = arr<T>(bb, Mcap);
            B.xd = nullptr;
        };
        const uint8_t* dineq = arr<uint8_t>(ineq, Mcap);
        iter(D, double{}, dx, dy, dax, daty, dxsum, dysum, dxlast, dylast, dc, dl, du, db);
        D.ineq = dineq;
        if (mixed) {
            V.fA = arr<float>(fA, Zcap);
            V.fAT = arr<float>(fAT, Zcap);
            iter(F, float{}, fx, fy, fax, faty, fxsum, fysum, fxlast, fylast, fc, fl, fu, fb);
            F.xd = arr<float>(fxd, 2 * Ncap);
            F.ineq = dineq;
        }
        K.c = arr<double>(cu, Ncap); K.l = arr<double>(lu, Ncap); K.u = arr<double>(uu, Ncap);
        K.rL = arr<double>(rLu, Mcap); K.rU = arr<double>(rUu, Mcap);
        K.Dr = arr<double>(Dr, Mcap); K.Dc = arr<double>(Dc, Ncap);
        K.xu = arr<double>(xu, Ncap); K.xu2 = arr<double>(xu2, Ncap); K.xavg = arr<double>(xavg, Ncap);
        K.aty_avg = arr<double>(atyavg, Ncap); K.best_x = arr<double>(bestx, Ncap);
        K.yu = arr<double>(yu, Mcap); K.yu2 = arr<double>(yu2, Mcap); K.yavg = arr<double>(yavg, Mcap);
        K.ax_avg = arr<double>(axavg, Mcap); K.best_y = arr<double>(besty, Mcap);
        arr<DevState>(dstate, 2 * (int64_t)Scap);
        arr<double>(ddx2, Scap);
        BATCH_HIP(hipHostMalloc((void**)&status_h, 2 * sizeof(double), hipHostMallocCoherent | hipHostMallocMapped));
        status_h[0] = status_h[1] = 0.0;
        BATCH_HIP(hipHostGetDevicePointer((void**)&status_d, status_h, 0));
        BATCH_HIP(hipHostMalloc((void**)&done_h, sizeof(int32_t) * Scap, hipHostMallocCoherent | hipHostMallocMapped));
        std::memset(done_h, 0, sizeof(int32_t) * Scap);
        BATCH_HIP(hipHostGetDevicePointer((void**)&done_d, done_h, 0));
        arr<BatchInst>(dinst, Scap);
        arr<int32_t>(dids, Scap);
        arr<DevState>(dstage, Scap);
        arr<int32_t>(dslots, Scap);
        set_config(b, c);
    }

    // Lanes rule and per-instance block cap (only while no instance runs).
    // Partial-sum and table storage: the resident instances' block counts
    // plus a full cap for every free slot.
    void set_config(double b, int c) {
        bias = b;
        cap = c;
        int64_t need_n = 0, need_m = 0, used = 0;
        for (int k = 0; k < Scap; ++k) {
            if (inst[k].n == 0) continue;
            shape_rule(k);
            need_n += inst[k].bn;
            need_m += inst[k].bm;
            ++used;
        }
        need_n += (Scap - used) * (int64_t)c;
        need_m += (Scap - used) * (int64_t)c;
        if (need_n + need_m > (int64_t)INT32_MAX / 16) throw ArgError("too many blocks for one launch");
        double* p = dparts.alloc<double>((size_t)(10 * need_n + 9 * need_m));
        part_n = p; p += need_n;
        part_kn = p; p += 8 * need_n;
        part_rn = p; p += need_n;
        part_m = p; p += 2 * need_m;
        part_km = p; p += 6 * need_m;
        part_rm = p;
        dtabn.alloc<BatchInst>(need_n);
        dtabm.alloc<BatchInst>(need_m);
    }

    // Lanes per row, staged or not, and block counts of instance k: functions
    // of its shape and the configuration only.
    void shape_rule(int k) {
        BatchInst& I = inst[k];
        I.LA = lanes_for(meanA[k], bias);
        I.LAT = lanes_for(meanAT[k], bias);
        I.sn = staged_side(I.n, I.LAT);
        I.sm = staged_side(I.m, I.LA);
        I.bn = I.sn ? tiles_for(I.n, cap) : blocks_for(I.n, I.LAT, cap);
        I.bm = I.sm ? tiles_for(I.m, cap) : blocks_for(I.m, I.LA, cap);
    }

    bool fits(const Prepared& p) const {
        return p.n + 1 <= Ncap && p.m + 1 <= Mcap && p.nnz <= Zcap;
    }

    // Takes ranges and a slot for p and uploads it; -1 when it does not fit now.
    int add(const Prepared& p) {
        if (free_slots.empty()) return -1;
        int64_t co = 0, ro = 0, za = 0;
        if (!an.alloc(p.n + 1, co)) return -1;
        if (!am.alloc(p.m + 1, ro)) { an.release(co, p.n + 1); return -1; }
        if (p.nnz > 0 && !az.alloc(p.nnz, za)) { an.release(co, p.n + 1); am.release(ro, p.m + 1); return -1; }
        const int k = free_slots.back();
        free_slots.pop_back();
        BatchInst& I = inst[k];
        I = BatchInst{};
        I.m = p.m; I.n = p.n;
        I.ro = ro; I.co = co; I.pa = ro; I.pat = co; I.za = za;
        I.nonneg = 1;
        for (int32_t j = 0; j < p.n && I.nonneg; ++j)
            I.nonneg = p.l[j] == 0.0 && !std::signbit(p.l[j]) && p.u[j] == std::numeric_limits<double>::infinity();
        nnz_slot[k] = p.nnz;
        meanA[k] = p.meanA;
        meanAT[k] = p.meanAT;
        shape_rule(k);
        init[k] = p.init;
        obj_offset[k] = p.init.obj_offset;
        const hipStream_t st = upload_stream ? upload_stream : stream;
        auto put = [st](auto* dst, const auto& v) {
            if (!v.empty())
                BATCH_HIP(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice, st));
        };
        try {
            upload(p, ro, co, za, put);
        } catch (...) {
            (void)hipStreamSynchronize(st);
            remove(k);   // ranges and slot back; the caller keeps the request
            throw;
        }
        return k;
    }

    // Copies p's arrays into the arenas at its ranges.
    template <typename Put>
    void upload(const Prepared& p, int64_t ro, int64_t co, int64_t za, Put&& put) {
        hipStream_t st = upload_stream ? upload_stream : stream;
        put(ptrA.get<int32_t>() + ro, p.ptrA);
        put(ptrAT.get<int32_t>() + co, p.ptrAT);
        put(idxA.get<int32_t>() + za, p.idxA);
        put(idxAT.get<int32_t>() + za, p.idxAT);
        put(dA.get<double>() + za, p.vA);
        put(dAT.get<double>() + za, p.vAT);
        put(uA.get<double>() + za, p.vuA);
        put(uAT.get<double>() + za, p.vuAT);
        put(dc.get<double>() + co, p.c);
        put(dl.get<double>() + co, p.l);
        put(du.get<double>() + co, p.u);
        put(db.get<double>() + ro, p.b);
        put(ineq.get<uint8_t>() + ro, p.ineq);
        put(cu.get<double>() + co, p.cu);
        put(lu.get<double>() + co, p.lu);
        put(uu.get<double>() + co, p.uu);
        put(Dc.get<double>() + co, p.Dc);
        put(rLu.get<double>() + ro, p.rLu);
        put(rUu.get<double>() + ro, p.rUu);
        put(Dr.get<double>() + ro, p.Dr);
        if (mixed) {
            put(fA.get<float>() + za, p.fA);
            put(fAT.get<float>() + za, p.fAT);
            put(fc.get<float>() + co, p.fc);
            put(fl.get<float>() + co, p.fl);
            put(fu.get<float>() + co, p.fu);
            put(fb.get<float>() + ro, p.fb);
        }
        BATCH_HIP(hipStreamSynchronize(st));   // p's vectors may go away after this
    }

    void remove(int k) {
        BatchInst& I = inst[k];
        an.release(I.co, I.n + 1);
        am.release(I.ro, I.m + 1);
        if (nnz_slot[k] > 0) az.release(I.za, nnz_slot[k]);
        nnz_slot[k] = 0;
        I = BatchInst{};
        free_slots.push_back(k);
    }

    // Lays the instances `live` out over the launches; returns the geometry.
    BatchGeo layout(const std::vector<int32_t>& live) {
        std::vector<BatchInst> tn, tm;   // per block: a copy of its instance's description
        int32_t on = 0, om = 0;
        for (int32_t k : live) {
            BatchInst& I = inst[k];
            I.on = on;
            I.om = om;
            I.id = k;
            tn.insert(tn.end(), I.bn, I);
            tm.insert(tm.end(), I.bm, I);
            on += I.bn;
            om += I.bm;
        }
        hipStream_t st = stream;
        BATCH_HIP(hipMemcpyAsync(dinst.get<BatchInst>(), inst.data(), sizeof(BatchInst) * Scap,
                                 hipMemcpyHostToDevice, st));
        BATCH_HIP(hipMemcpyAsync(dtabn.get<BatchInst>(), tn.data(), sizeof(BatchInst) * tn.size(),
                                 hipMemcpyHostToDevice, st));
        BATCH_HIP(hipMemcpyAsync(dtabm.get<BatchInst>(), tm.data(), sizeof(BatchInst) * tm.size(),
                                 hipMemcpyHostToDevice, st));
        BATCH_HIP(hipMemcpyAsync(dids.get<int32_t>(), live.data(), sizeof(int32_t) * live.size(),
                                 hipMemcpyHostToDevice, st));
        BATCH_HIP(hipStreamSynchronize(st));   // the host vectors go out of scope
        BatchGeo g;
        g.inst = dinst.get<BatchInst>();
        g.blk_n = dtabn.get<BatchInst>();
        g.blk_m = dtabm.get<BatchInst>();
        g.ids = dids.get<int32_t>();
        g.BN = on;
        g.BM = om;
        g.ptrA = ptrA.get<int32_t>();
        g.idxA = idxA.get<int32_t>();
        g.ptrAT = ptrAT.get<int32_t>();
        g.idxAT = idxAT.get<int32_t>();
        return g;
    }

    // Cold start of `fresh` (initial states with this iteration limit,
    // starting iterates).  Lays out only `fresh`: the caller lays out the
    // running set afterwards.
    void cold_start(const std::vector<int32_t>& fresh, int max_iter) {
        if (fresh.empty()) return;
        std::vector<DevState> st0(fresh.size());
        for (int32_t k : fresh) done_h[k] = 0;
        for (size_t r = 0; r < fresh.size(); ++r) {
            DevState s = init[fresh[r]];
            s.max_iter = max_iter;
            s.active = max_iter > 0 ? 1 : 0;
            s.need_check = max_iter > 0 ? 0 : 1;
            st0[r] = s;
        }
        hipStream_t st = stream;
        BATCH_HIP(hipMemcpyAsync(dstage.get<DevState>(), st0.data(), sizeof(DevState) * st0.size(),
                                 hipMemcpyHostToDevice, st));
        BATCH_HIP(hipMemcpyAsync(dslots.get<int32_t>(), fresh.data(), sizeof(int32_t) * fresh.size(),
                                 hipMemcpyHostToDevice, st));
        const int cnt = (int)fresh.size();
        hipLaunchKernelGGL(kb_put_states, dim3((cnt + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                           dstate.get<DevState>(), dstage.get<DevState>(), dslots.get<int32_t>(), cnt);
        const BatchGeo g = layout(fresh);
        if (mixed) {
            hipLaunchKernelGGL(kb_cold<float>, dim3(g.BN + g.BM), dim3(kBlock), 0, st, g, F);
        } else {
            hipLaunchKernelGGL(kb_cold<double>, dim3(g.BN + g.BM), dim3(kBlock), 0, st, g, D);
            hipLaunchKernelGGL(kb_cold_ax<double>, dim3(g.BM), dim3(kBlock), 0, st, g, V, D);
        }
        BATCH_HIP(hipGetLastError());
        BATCH_HIP(hipStreamSynchronize(st));
    }

    // marks (diagnostics only, never while capturing): 2G + 2 events, one
    // before each attempt launch, one after the last, one after the checks.
    void enqueue_replay(const BatchGeo& g, int live, int G, hipEvent_t* marks = nullptr) {
        hipStream_t st = stream;
        DevState* S = dstate.get<DevState>();
        double* dx2 = ddx2.get<double>();
        for (int a = 0; a < G; ++a) {
            const int p = a & 1;
            if (marks) BATCH_HIP(hipEventRecord(marks[2 * a], st));
            hipLaunchKernelGGL(kb_primal, dim3(g.BN), dim3(kBlock), 0, st, S, p, g, V, F, D, part_m, dx2, part_n);
            if (marks) BATCH_HIP(hipEventRecord(marks[2 * a + 1], st));
            hipLaunchKernelGGL(kb_dual, dim3(g.BM), dim3(kBlock), 0, st, S, p ^ 1, g, V, F, D, part_n, dx2, part_m);
        }
        if (marks) BATCH_HIP(hipEventRecord(marks[2 * G], st));
        hipLaunchKernelGGL(kb_check_prep, dim3(g.BN + g.BM), dim3(kBlock), 0, st, S, g, F, D, K, part_m, dx2);
        hipLaunchKernelGGL(kb_check_kkt, dim3(g.BN + g.BM), dim3(kBlock), 0, st, S, g, V, K, part_kn, part_km);
        hipLaunchKernelGGL(kb_kkt_fin, dim3(live), dim3(kBlock), 0, st, S, g, part_kn, part_km);
        hipLaunchKernelGGL(kb_check_restart, dim3(g.BN + g.BM), dim3(kBlock), 0, st, S, g, F, D, K, part_rn, part_rm);
        if (mixed)
            hipLaunchKernelGGL(kb_switch_products, dim3(g.BN + g.BM), dim3(kBlock), 0, st, S, g, V, D);
        hipLaunchKernelGGL(kb_restart_fin, dim3(live), dim3(kBlock), 0, st, S, g, part_rn, part_rm, done_d);
        hipLaunchKernelGGL(kb_status, dim3(1), dim3(kBlock), 0, st, S, g, live, status_d);
        if (marks) BATCH_HIP(hipEventRecord(marks[2 * G + 1], st));
        BATCH_HIP(hipGetLastError());
    }

    // One replay launched kernel by kernel with events between the launches
    // (verbose >= 3): prints where a replay's time goes.  It is a real
    // replay: the instances advance exactly as under the graph.
    void profile_replay(const BatchGeo& g, int live, int G) {
        std::vector<hipEvent_t> ev(2 * G + 2, nullptr);
        for (auto& e : ev) BATCH_HIP(hipEventCreate(&e));
        enqueue_replay(g, live, G, ev.data());
        BATCH_HIP(hipStreamSynchronize(stream));
        float prim = 0.f, dual = 0.f, check = 0.f, t = 0.f;
        for (int a = 0; a < G; ++a) {
            BATCH_HIP(hipEventElapsedTime(&t, ev[2 * a], ev[2 * a + 1]));
            prim += t;
            BATCH_HIP(hipEventElapsedTime(&t, ev[2 * a + 1], ev[2 * a + 2]));
            dual += t;
        }
        BATCH_HIP(hipEventElapsedTime(&check, ev[2 * G], ev[2 * G + 1]));
        for (auto& e : ev) (void)hipEventDestroy(e);
        std::fprintf(stderr, "[suprema] batch profile: %d running, blocks %d + %d: primal %.1f us, dual %.1f us per attempt; check group %.1f us\n",
                     live, g.BN, g.BM, 1e3 * prim / G, 1e3 * dual / G, 1e3 * check);
    }

    hipGraphExec_t capture(const BatchGeo& g, int live, int G) {
        hipGraph_t gr = nullptr;
        hipGraphExec_t exec = nullptr;
        BATCH_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            for (int u = 0; u < units; ++u) enqueue_replay(g, live, G);
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

    int graph_attempts() const {
        int G = opt.graph_attempts != 0 ? std::abs(opt.graph_attempts) : 16;
        return G + (G & 1);   // the state double buffer needs an even number per replay
    }
    // graph_attempts < 0: each replay's kernel sequence is launched on the
    // stream directly instead of from a captured graph (the same kernels in
    // the same order: the same bits; for tools that trace kernel launches)
    bool direct() const { return opt.graph_attempts < 0; }

    // One replay: the captured graph, or the same launches made directly.
    int units = 1;   // units of (G attempts + check group) per replay or poll
    void launch(hipGraphExec_t exec, const BatchGeo& g, int live, int G) {
        if (direct()) for (int u = 0; u < units; ++u) enqueue_replay(g, live, G);
        else BATCH_HIP(hipGraphLaunch(exec, stream));
    }

    void read_states(std::vector<DevState>& h) {
        h.resize(2 * (size_t)Scap);
        BATCH_HIP(hipMemcpy(h.data(), dstate.get<DevState>(), sizeof(DevState) * h.size(),
                            hipMemcpyDeviceToHost));
    }

    // The point and information of a finished instance.
    void result(int k, const DevState& s, double* x_out, double* y_out, supr_pdlp_info& inf) {
        const BatchInst& I = inst[k];
        std::memset(&inf, 0, sizeof(inf));
        inf.iterations = s.attempts;
        inf.restarts = s.restarts;
        inf.converged = s.done == 1 ? 1 : 0;
        inf.has_point = s.has_best ? 1 : 0;
        inf.eps_P = s.best_eps[0]; inf.eps_D = s.best_eps[1]; inf.eps_G = s.best_eps[2];
        inf.iterations_fp32 = mixed ? (s.phase == 0 ? s.attempts : s.attempts_fp32) : 0;
        inf.checks = s.checks;
        inf.lanes_A = I.LA; inf.lanes_AT = I.LAT;
        double obj = obj_offset[k];
        if (s.has_best) {
            // the returned point; its objective c'x + offset (unscaled form)
            // was formed by the KKT check that kept it
            obj = s.best_obj;
            if (x_out) BATCH_HIP(hipMemcpy(x_out, K.best_x + I.co, sizeof(double) * I.n, hipMemcpyDeviceToHost));
            if (y_out) BATCH_HIP(hipMemcpy(y_out, K.best_y + I.ro, sizeof(double) * I.m, hipMemcpyDeviceToHost));
        }
        inf.objective = obj;
    }
};

// ------------------------------------------------------------ the batch --

// Per-instance block cap of the batched engine: large enough that the
// largest instances, which finish last, are not starved of blocks.
constexpr int kBatchCap = 512;

struct Batch {
    Pool pool;
    int count = 0;
    std::vector<int32_t> slot;   // instance -> pool slot

    void build(int32_t cnt, const supr_lp* U, const supr_lp* S, const double* const* Drs,
               const double* const* Dcs, const supr_pdlp_options& o) {
        count = cnt;
        const bool mixed = o.mode != SUPR_MODE_FAST_FP64;
        int64_t N = 0, M = 0, Z = 0;
        for (int k = 0; k < cnt; ++k) { N += S[k].n + 1; M += S[k].m + 1; Z += S[k].nnz; }
        pool.create(N, M, Z, cnt, o, 1.0, kBatchCap);
        slot.resize(cnt);
        for (int k = 0; k < cnt; ++k) {
            slot[k] = pool.add(prepare(U[k], S[k], Drs[k], Dcs[k], o, mixed));
            if (slot[k] < 0) throw std::runtime_error("batch layout: an instance does not fit its own pool");
        }
        pool.set_config(pool.bias, pool.cap);   // size the partial sums exactly
    }

    // Solves `ids` from the cold start; states[r] is the final state of ids[r].
    double run(const std::vector<int32_t>& ids, int max_iter, std::vector<DevState>& states) {
        std::vector<int32_t> live;
        {
            std::vector<char> seen(count, 0);
            for (int32_t k : ids) {
                if (k < 0 || k >= count) throw ArgError("instance id out of range");
                if (seen[k]) throw ArgError("instance id listed twice");
                seen[k] = 1;
                live.push_back(slot[k]);
            }
        }
        hipStream_t st = pool.stream;
        const auto t0 = Clock::now();
        pool.cold_start(live, max_iter);
        BatchGeo g = pool.layout(live);
        const int G = pool.graph_attempts();
        // auto: 4 units (G attempts + the check group) per replay or poll,
        // fewer graph-launch boundaries (the 512-LP blending batch: 6.27 ->
        // 6.19 s, the same bits); an explicit graph_attempts: one unit
        pool.units = pool.opt.graph_attempts == 0 ? 4 : 1;
        hipGraphExec_t exec = nullptr;
        constexpr int R = 4;
        hipEvent_t ev[R] = {};
        std::vector<DevState> h;
        auto cleanup = [&]() {
            (void)hipStreamSynchronize(st);
            for (auto& e : ev) if (e) { (void)hipEventDestroy(e); e = nullptr; }
            if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
        };
        try {
            for (auto& e : ev) BATCH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            if (pool.opt.verbose >= 3) pool.profile_replay(g, (int)live.size(), G);
            if (!pool.direct()) exec = pool.capture(g, (int)live.size(), G);
            double last_progress = -1.0;
            int idle = 0, replays = 0, inflight = 0;
            size_t live_at_capture = live.size();
            while (true) {
                if (inflight == R) {
                    const int s = (replays - inflight) % R;
                    BATCH_HIP(hipEventSynchronize(ev[s]));
                    --inflight;
                    // the status the replays' last kernel wrote (this replay's or a
                    // newer one's: both only move forward)
                    std::atomic_thread_fence(std::memory_order_acquire);
                    const double running = ((volatile double*)pool.status_h)[0];
                    const double progress = ((volatile double*)pool.status_h)[1];
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
                        pool.read_states(h);
                        std::vector<int32_t> still;
                        for (int32_t k : live) if (!h[2 * k].done) still.push_back(k);
                        if (still.empty()) break;
                        if (pool.opt.verbose)
                            std::fprintf(stderr, "[suprema] batch: replay %d, %zu of %zu instances running, layout compacted\n",
                                         replays, still.size(), live.size());
                        live.swap(still);
                        live_at_capture = live.size();
                        if (exec) (void)hipGraphExecDestroy(exec);
                        exec = nullptr;
                        g = pool.layout(live);
                        if (pool.opt.verbose >= 3) pool.profile_replay(g, (int)live.size(), G);
                        if (!pool.direct()) exec = pool.capture(g, (int)live.size(), G);
                    }
                }
                const int s = replays % R;
                pool.launch(exec, g, (int)live.size(), G);
                BATCH_HIP(hipEventRecord(ev[s], st));
                ++replays;
                ++inflight;
            }
            BATCH_HIP(hipStreamSynchronize(st));
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
        const double secs = std::chrono::duration<double>(Clock::now() - t0).count();
        pool.read_states(h);
        states.resize(ids.size());
        for (size_t r = 0; r < ids.size(); ++r) states[r] = h[2 * slot[ids[r]]];
        return secs;
    }

    // Times every lanes-rule candidate on a short run of the whole batch
    // and keeps the fastest.  The block cap stays at its setting: it only
    // limits the largest instances, which decide the tail of a solve that a
    // short probe does not reach.
    void tune(int probe_iters) {
        std::vector<int32_t> all(count);
        std::iota(all.begin(), all.end(), 0);
        std::vector<DevState> scratch_states;
        double best = std::numeric_limits<double>::infinity(), best_bias = 1.0;
        const int c = pool.cap;
        for (double b : {0.5, 1.0, 2.0}) {
            pool.set_config(b, c);
            run(all, probe_iters, scratch_states);            // warm
            const double t = run(all, probe_iters, scratch_states);
            if (t < best) { best = t; best_bias = b; }
        }
        pool.set_config(best_bias, c);
    }
};

// ----------------------------------------------------------- the server --

struct Request {
    int64_t ticket = 0;
    Prepared prep;
    Clock::time_point submitted;
};

struct Finished {
    int32_t m = 0, n = 0;
    std::vector<double> x, y;
    supr_pdlp_info info{};
};

struct Server {
    Pool pool;
    supr_server_options so{};
    std::mutex mu;                      // guards queue, next_ticket, done
    std::deque<std::unique_ptr<Request>> queue;
    int64_t next_ticket = 1;
    std::unordered_map<int64_t, Finished> done;
    // owned by the stepping thread
    std::vector<int32_t> live;          // running slots
    std::unordered_map<int32_t, std::unique_ptr<Request>> running;   // slot -> request
    std::unordered_map<int32_t, Clock::time_point> admitted;
    hipGraphExec_t exec = nullptr;
    BatchGeo geo{};                     // the running set's layout (direct launches)
    bool recapture = false;             // the running set changed since the graph was captured
    int32_t finished_total = 0;
    // steps whose replays may still run (at most two), with their end events
    hipEvent_t ev[2] = {nullptr, nullptr};
    int inflight = 0;
    int64_t steps = 0;

    ~Server() {
        if (pool.stream) (void)hipStreamSynchronize(pool.stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        for (auto& e : ev) if (e) (void)hipEventDestroy(e);
    }

    // Waits for every launched replay (before the running set changes: the
    // replays in flight read its layout, slots and ranges).
    void drain() {
        BATCH_HIP(hipStreamSynchronize(pool.stream));
        inflight = 0;
    }

    // Undoes the admission of `fresh` (not yet running): slots back to the
    // pool, requests back to the head of the queue in their order.
    void unadmit(std::vector<int32_t>& fresh) {
        std::lock_guard<std::mutex> lock(mu);
        for (auto k = fresh.rbegin(); k != fresh.rend(); ++k) {
            auto it = running.find(*k);
            if (it != running.end()) {
                queue.push_front(std::move(it->second));
                running.erase(it);
            }
            admitted.erase(*k);
            pool.remove(*k);
        }
        fresh.clear();
    }

    void create(const supr_server_options& o) {
        so = o;
        const int cap = o.max_blocks > 0 ? o.max_blocks : 1024;
        const double bias = o.lanes_bias > 0 ? o.lanes_bias : 1.0;
        pool.create(o.max_cols, o.max_rows, o.max_nnz, o.max_instances, o.pdlp, bias, cap);
        BATCH_HIP(hipStreamCreateWithFlags(&pool.upload_stream, hipStreamNonBlocking));
        for (auto& e : ev) BATCH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }

    int64_t submit(const supr_lp& U, const supr_lp& S, const double* Dr, const double* Dc) {
        auto r = std::make_unique<Request>();
        r->prep = prepare(U, S, Dr, Dc, pool.opt, pool.mixed);   // caller's thread, no lock held
        r->submitted = Clock::now();
        if (!pool.fits(r->prep)) throw ArgError("request exceeds the server's column, row or nonzero capacity");
        std::lock_guard<std::mutex> lock(mu);
        r->ticket = next_ticket++;
        const int64_t t = r->ticket;
        queue.push_back(std::move(r));
        return t;
    }

    // Retires finished instances, admits queued requests that fit (first
    // come, first served: a request that does not fit waits at the head),
    // then runs the configured number of graph replays.  State stays
    // consistent when a step throws: a slot leaves `live`, `running` and the
    // pool together, after its result is stored; a request whose admission
    // fails goes back to the head of the queue with its host data; a failed
    // capture is retried at the next step.
    // One step.  Admitted requests are uploaded (upload stream) while the
    // running set's replays still run; the running set changes -- finished
    // requests retired, admitted ones started -- only after the replays in
    // flight have drained; then this step's replays are launched, with at most
    // two steps' replays in flight (the host does not wait for them unless it
    // must).  Finished requests are seen through the host-coherent done flags,
    // which only move forward.
    void step(int32_t* n_running, int32_t* n_queued, int32_t* n_finished) {
        std::atomic_thread_fence(std::memory_order_acquire);
        bool any_done = false;
        for (int32_t k : live)
            if (((volatile int32_t*)pool.done_h)[k]) { any_done = true; break; }
        std::vector<int32_t> fresh;
        while (true) {
            std::unique_ptr<Request> r;
            {
                std::lock_guard<std::mutex> lock(mu);
                if (queue.empty()) break;
                r = std::move(queue.front());
                queue.pop_front();
            }
            int k = -1;
            try {
                k = pool.add(r->prep);   // free ranges and slot: no replay in flight reads them
            } catch (...) {
                std::lock_guard<std::mutex> lock(mu);
                queue.push_front(std::move(r));
                drain();
                unadmit(fresh);
                throw;
            }
            if (k < 0) {
                std::lock_guard<std::mutex> lock(mu);
                queue.push_front(std::move(r));
                break;
            }
            admitted[k] = Clock::now();
            running[k] = std::move(r);
            fresh.push_back(k);
        }
        if (any_done || !fresh.empty()) drain();
        if (any_done) {
            for (size_t q = 0; q < live.size();) {
                const int32_t k = live[q];
                if (!((volatile int32_t*)pool.done_h)[k]) { ++q; continue; }
                auto it = running.find(k);
                if (it == running.end()) throw std::logic_error("server: running slot without a request");
                DevState s;
                BATCH_HIP(hipMemcpy(&s, pool.dstate.get<DevState>() + 2 * k, sizeof(DevState),
                                    hipMemcpyDeviceToHost));
                Finished f;
                f.m = pool.inst[k].m;
                f.n = pool.inst[k].n;
                f.x.resize(f.n);
                f.y.resize(f.m);
                pool.result(k, s, f.x.data(), f.y.data(), f.info);
                const auto now = Clock::now();
                f.info.seconds_solve = std::chrono::duration<double>(now - admitted[k]).count();
                f.info.seconds_total = std::chrono::duration<double>(now - it->second->submitted).count();
                f.info.seconds_setup = f.info.seconds_total - f.info.seconds_solve;
                const int64_t t = it->second->ticket;
                {
                    std::lock_guard<std::mutex> lock(mu);
                    done.emplace(t, std::move(f));
                }
                live.erase(live.begin() + (std::ptrdiff_t)q);
                running.erase(it);
                admitted.erase(k);
                pool.remove(k);
                ++finished_total;
                recapture = true;
            }
        }
        if (!fresh.empty()) {
            try {
                pool.cold_start(fresh, pool.opt.max_iterations);
            } catch (...) {
                unadmit(fresh);
                throw;
            }
            for (int32_t k : fresh) running[k]->prep = Prepared{};   // uploaded; free the host copy
            live.insert(live.end(), fresh.begin(), fresh.end());
            recapture = true;
        }
        if (recapture) {
            // the set changed after a drain: nothing in flight reads the old layout
            if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
            if (!live.empty()) {
                geo = pool.layout(live);
                if (!pool.direct()) exec = pool.capture(geo, (int)live.size(), pool.graph_attempts());
                if (pool.opt.verbose)
                    std::fprintf(stderr, "[suprema] server: %zu running, graph captured\n", live.size());
            }
            recapture = false;
        }
        if (!live.empty()) {
            const int slot = (int)(steps % 2);
            if (inflight == 2) {
                BATCH_HIP(hipEventSynchronize(ev[slot]));
                --inflight;
            }
            const int reps = so.replays_per_step > 0 ? so.replays_per_step : 1;
            for (int q = 0; q < reps; ++q) pool.launch(exec, geo, (int)live.size(), pool.graph_attempts());
            BATCH_HIP(hipEventRecord(ev[slot], pool.stream));
            ++inflight;
            ++steps;
        }
        *n_running = (int32_t)live.size();
        {
            std::lock_guard<std::mutex> lock(mu);
            *n_queued = (int32_t)queue.size();
        }
        *n_finished = finished_total;
    }
};

}  // namespace supr

struct supr_batch {
    supr::Batch b;
};

struct supr_server {
    supr::Server s;
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
        if (!check_instance(unscaled_forms[k], scaled_forms[k], Dr[k], Dc[k],
                            "supr_batch_create: instance " + std::to_string(k)))
            return SUPR_EINVAL;
        zsum += scaled_forms[k].nnz;
    }
    if (zsum > INT32_MAX) { set_error("supr_batch_create: total nnz exceeds int32"); return SUPR_EINVAL; }
    int devs = 0;
    if (hipGetDeviceCount(&devs) != hipSuccess || devs < 1) { set_error("no HIP device visible"); return SUPR_ENODEV; }
    try {
        BATCH_HIP(hipSetDevice(opt->device));
        auto h = std::make_unique<supr_batch>();
        h->b.pool.device = opt->device;
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
        BATCH_HIP(hipSetDevice(batch->b.pool.device));
        batch->b.pool.set_config(lanes_bias, max_blocks);
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
    *lanes_bias = batch->b.pool.bias;
    *max_blocks = batch->b.pool.cap;
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
        BATCH_HIP(hipSetDevice(b.pool.device));
        std::vector<int32_t> list;
        if (ids) list.assign(ids, ids + nids);
        else { list.resize(b.count); std::iota(list.begin(), list.end(), 0); }
        std::vector<DevState> st;
        const double secs = b.run(list, b.pool.opt.max_iterations, st);
        if (seconds) *seconds = secs;
        for (size_t r = 0; r < list.size(); ++r) {
            b.pool.result(b.slot[list[r]], st[r], x_out ? x_out[r] : nullptr, y_out ? y_out[r] : nullptr,
                          infos[r]);
            infos[r].seconds_solve = secs;
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
    (void)hipSetDevice(batch->b.pool.device);
    delete batch;
}

void supr_server_default_options(supr_server_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    supr_pdlp_default_options(&o->pdlp);
    o->max_cols = (int64_t)1 << 24;
    o->max_rows = (int64_t)1 << 23;
    o->max_nnz = (int64_t)1 << 27;
    o->max_instances = 256;
    o->max_blocks = 1024;
    o->lanes_bias = 1.0;
    o->replays_per_step = 1;
}

int64_t supr_server_request_bytes(int32_t m, int32_t n, int32_t nnz, int32_t mode) {
    return supr::footprint_bytes(m, n, nnz, mode != SUPR_MODE_FAST_FP64);
}

int supr_server_create(const supr_server_options* opt, supr_server** out) {
    using namespace supr;
    if (!opt || !out) { set_error("supr_server_create: null argument"); return SUPR_EINVAL; }
    *out = nullptr;
    const supr_pdlp_options& p = opt->pdlp;
    if (p.mode == SUPR_MODE_PLAIN_FP64 || p.max_iterations < 0 || p.check_every < 1 || !(p.tol >= 0) ||
        opt->max_cols < 2 || opt->max_rows < 2 || opt->max_nnz < 1 || opt->max_instances < 1 ||
        opt->max_blocks < 0 || opt->max_blocks > kMaxBlocks) {
        set_error("supr_server_create: bad options (device-resident mode, positive capacities, max_blocks <= 2048)");
        return SUPR_EINVAL;
    }
    int devs = 0;
    if (hipGetDeviceCount(&devs) != hipSuccess || devs < 1) { set_error("no HIP device visible"); return SUPR_ENODEV; }
    try {
        BATCH_HIP(hipSetDevice(p.device));
        auto h = std::make_unique<supr_server>();
        h->s.pool.device = p.device;
        h->s.create(*opt);
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

int supr_server_device_bytes(const supr_server* server, int64_t* bytes) {
    if (!server || !bytes) { supr::set_error("null argument"); return SUPR_EINVAL; }
    *bytes = server->s.pool.device_bytes;
    return SUPR_OK;
}

int supr_server_submit(supr_server* server, const supr_lp* unscaled_form, const supr_lp* scaled_form,
                       const double* Dr, const double* Dc, int64_t* ticket) {
    using namespace supr;
    if (!server || !unscaled_form || !scaled_form || !ticket) {
        set_error("supr_server_submit: null argument");
        return SUPR_EINVAL;
    }
    if (!check_instance(*unscaled_form, *scaled_form, Dr, Dc, "supr_server_submit")) return SUPR_EINVAL;
    try {
        *ticket = server->s.submit(*unscaled_form, *scaled_form, Dr, Dc);
        return SUPR_OK;
    } catch (const ArgError& e) {
        set_error(e.what());
        return SUPR_EINVAL;
    } catch (const std::bad_alloc&) {
        set_error("out of host memory");
        return SUPR_ENOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SUPR_EINTERNAL;
    }
}

int supr_server_step(supr_server* server, int32_t* running, int32_t* queued, int32_t* finished) {
    using namespace supr;
    if (!server || !running || !queued || !finished) { set_error("supr_server_step: null argument"); return SUPR_EINVAL; }
    try {
        BATCH_HIP(hipSetDevice(server->s.pool.device));
        server->s.step(running, queued, finished);
        return SUPR_OK;
    } catch (const HipError& e) {
        set_error(e.what());
        return SUPR_EDEVICE;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SUPR_EINTERNAL;
    }
}

int supr_server_result(supr_server* server, int64_t ticket, int32_t* m, int32_t* n, double* x_out,
                       double* y_out, supr_pdlp_info* info, int32_t* ready) {
    using namespace supr;
    if (!server || !ready) { set_error("supr_server_result: null argument"); return SUPR_EINVAL; }
    std::lock_guard<std::mutex> lock(server->s.mu);
    auto it = server->s.done.find(ticket);
    *ready = it != server->s.done.end() ? 1 : 0;
    if (!*ready) return SUPR_OK;
    const Finished& f = it->second;
    if (m) *m = f.m;
    if (n) *n = f.n;
    if (!x_out && !y_out && !info) return SUPR_OK;   // size query: the result stays
    if (x_out) std::copy(f.x.begin(), f.x.end(), x_out);
    if (y_out) std::copy(f.y.begin(), f.y.end(), y_out);
    if (info) *info = f.info;
    server->s.done.erase(it);
    return SUPR_OK;
}

void supr_server_destroy(supr_server* server) {
    if (!server) return;
    (void)hipSetDevice(server->s.pool.device);
    delete server;
}

}  // extern "C"
