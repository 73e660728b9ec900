// pdlp_engine.hip -- the MI355X restarted-PDHG engine behind
// supr_solve_pdhg_gpu (the reference's solve_pdhg_gpu,
// gpu/pdlp_gpu_solve.cu:261-267 + include/pdlp_algo.hpp:49-185).
//
// Modes (see suprema_pdlp.h):
//  * PLAIN_FP64 -- the reference GPU backend's design on this GPU: FP64
//    kernels, the host reads the step statistics after every attempt and
//    runs the KKT check (product verify on the host) every check_every
//    accepted steps.  It is the honest "plain FP64 GPU PDLP" baseline.
//  * FAST_FP64 / MIXED -- the device-resident loop: the algorithm's control
//    flow lives in a DevState block on the device, iterations and checks are
//    replayed from a captured hipGraph, the host only polls a 200-byte
//    state copy once per replay.  MIXED runs FP32 iterates / FP32 SpMV with
//    FP64 statistics, KKT and restart rule, then continues in FP64 from the
//    FP32 loop's candidate (the polish) until the tolerance is verified.
//
// Memory layout (per LP, all resident in HBM for the whole solve):
//   A    : CSR (row pointers, int32 column indices) or ELL for uniform short
//          rows; values as FP32 scaled, FP64 scaled, FP64 unscaled arrays in
//          the same order (one index stream serves all three).
//   A^T  : the same for the stored transpose (reference keeps both,
//          gpu/pdlp_gpu_solve.cu:194).
//   iterates: x, y, A x double-buffered (swap = flip of DevState::cur),
//          A^T y, ergodic sums, restart points; c, l, u, b, inequality flags.
//   KKT  : unscaled c, l, u, rL, rU, Dr, Dc; unscaled candidates, averages,
//          best point (FP64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <cstring>
#include <future>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "hip_util.hpp"
#include "host_lp.hpp"
#include "pdlp_kernels.hpp"
#include "prep.hpp"
#include "suprema_pdlp.h"

namespace supr {

// ------------------------------------------------------ device memory cache --

namespace {
struct CachedBlock {
    void* p;
    size_t bytes;
    int device;
};
struct DeviceCache {
    std::mutex mu;
    std::vector<CachedBlock> blocks;
    size_t held = 0;
};
// never destroyed: no HIP call may run while the process exits
DeviceCache& device_cache() {
    static DeviceCache* c = new DeviceCache();
    return *c;
}
size_t device_cache_cap() {
    static const size_t cap = [] {
        const char* e = std::getenv("SUPR_DEVICE_CACHE_MB");
        return (e ? (size_t)std::strtoull(e, nullptr, 10) : (size_t)4096) << 20;
    }();
    return cap;
}
void release_cached(int device) {   // device < 0: all devices
    std::vector<CachedBlock> drop;
    {
        DeviceCache& c = device_cache();
        std::lock_guard<std::mutex> g(c.mu);
        for (size_t k = 0; k < c.blocks.size();) {
            if (device < 0 || c.blocks[k].device == device) {
                c.held -= c.blocks[k].bytes;
                drop.push_back(c.blocks[k]);
                c.blocks[k] = c.blocks.back();
                c.blocks.pop_back();
            } else {
                ++k;
            }
        }
    }
    for (const CachedBlock& b : drop) (void)hipFree(b.p);
}
}  // namespace

void* device_alloc(size_t bytes, size_t* reserved, int* device) {
    int dev = 0;
    SUPR_HIP(hipGetDevice(&dev));
    {
        // the smallest cached block of this device that fits, if it is at
        // most twice the request
        DeviceCache& c = device_cache();
        std::lock_guard<std::mutex> g(c.mu);
        size_t best = c.blocks.size();
        for (size_t k = 0; k < c.blocks.size(); ++k) {
            const CachedBlock& b = c.blocks[k];
            if (b.device == dev && b.bytes >= bytes && b.bytes <= 2 * bytes &&
                (best == c.blocks.size() || b.bytes < c.blocks[best].bytes))
                best = k;
        }
        if (best < c.blocks.size()) {
            const CachedBlock b = c.blocks[best];
            c.blocks[best] = c.blocks.back();
            c.blocks.pop_back();
            c.held -= b.bytes;
            *reserved = b.bytes;
            *device = dev;
            return b.p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory) {   // give the cached blocks back and try again
        (void)hipGetLastError();
        release_cached(dev);
        e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess)
        throw HipFailure(std::string("HIP error: hipMalloc of ") + std::to_string(bytes) + " bytes -- " +
                         hipGetErrorString(e));
    *reserved = bytes;
    *device = dev;
    return p;
}

void device_free(void* p, size_t reserved, int device) {
    if (!p) return;
    int cur = -1;
    const bool same = hipGetDevice(&cur) == hipSuccess && cur == device;
    if (same && reserved <= device_cache_cap() && hipDeviceSynchronize() == hipSuccess) {
        DeviceCache& c = device_cache();
        std::lock_guard<std::mutex> g(c.mu);
        if (c.held + reserved <= device_cache_cap()) {
            c.blocks.push_back({p, reserved, device});
            c.held += reserved;
            return;
        }
    }
    (void)hipFree(p);
}

namespace {

// Row-loop grids are rounded up to a multiple of the XCD count (the kernels'
// XCD-aware block mapping needs it; extra blocks find no rows).
int grid_rows(int rows, int L) {
    // staged schedules (pdlp_device.hpp) work in tiles of kBlock rows
    const int per = (L >= 2 && rows >= kStageRows) ? kBlock : (L >= kBlock ? 1 : kBlock / L);
    long g = ((long)rows + per - 1) / per;
    g = std::max(1L, std::min<long>(g, kMaxBlocks));
    if (g >= kXcds) g = (g + kXcds - 1) / kXcds * kXcds;
    return (int)g;
}
int grid_elems(long len) {
    return (int)std::max(1L, std::min<long>((len + kBlock - 1) / kBlock, kMaxBlocks));
}
// The check group's grids: at most 512 blocks (two per CU), grid-stride over
// the columns and rows.  Most replays have no check due, and a launch whose
// blocks all return at once costs about 1 us per 512 blocks dispatched.
constexpr int kCheckBlocks = 512;
int grid_check_elems(long len) { return std::min(grid_elems(len), kCheckBlocks); }
int grid_check_rows(int rows, int L) {
    const int g = grid_rows(rows, L);
    return g > kCheckBlocks ? kCheckBlocks : g;
}

#define SUPR_LANES(L, ...)                                                        \
    switch (L) {                                                                  \
        case 1: { constexpr int LL = 1; __VA_ARGS__; } break;                     \
        case 2: { constexpr int LL = 2; __VA_ARGS__; } break;                     \
        case 4: { constexpr int LL = 4; __VA_ARGS__; } break;                     \
        case 8: { constexpr int LL = 8; __VA_ARGS__; } break;                     \
        case 16: { constexpr int LL = 16; __VA_ARGS__; } break;                   \
        case 32: { constexpr int LL = 32; __VA_ARGS__; } break;                   \
        case 64: { constexpr int LL = 64; __VA_ARGS__; } break;                   \
        case 256: { constexpr int LL = 256; __VA_ARGS__; } break;                 \
        default: throw BadInput("unsupported lanes-per-row " + std::to_string(L)); \
    }

bool valid_lanes(int L) {
    return L == 1 || L == 2 || L == 4 || L == 8 || L == 16 || L == 32 || L == 64 || L == kBlock;
}

// --------------------------------------------------------- device matrix --

struct DevMatrix {
    int rows = 0, cols = 0, nnz = 0, ell_w = 0, lanes = 1;
    DevBuf ptr, idx, idx16, v32, v64, vu;
    // ELL, FP64: the factors the device-resident loop forms the values from
    // (MatView::rsc / csc), set when they reproduce every stored value
    const double* rsc = nullptr;
    const double* csc = nullptr;
    template <typename V> MatView<V> view(const DevBuf& v) const {
        MatView<V> m;
        m.rows = rows;
        m.ell_w = ell_w;
        m.ptr = ptr.as<int32_t>();
        m.idx = idx.as<int32_t>();
        m.val = v.as<V>();
        m.idx16 = idx16.as<uint16_t>();
        return m;
    }
    MatView<float> f32() const { return view<float>(v32); }
    MatView<double> f64() const { return view<double>(v64); }
    // the device-resident loop's FP64 view (the plain mode reads the values)
    MatView<double> f64_iter() const {
        MatView<double> m = f64();
        m.rsc = rsc;
        m.csc = csc;
        return m;
    }
    MatView<double> unscaled() const { return view<double>(vu); }
};

constexpr int kMaxEllWidth = 16;

// ELL indices narrower than int32: when every index is below 65,536 (the
// gathered vector is that short: A^T of an LP with few rows), the ELL loops
// read a uint16 copy -- on the 1M-variable transportation LP's A^T that is
// 4 MB instead of 8 MB per primal launch.  The int32 copy stays for the host
// side and the layout checks.
bool narrow_ell(const DevMatrix& D) { return D.ell_w > 0 && D.cols > 0 && D.cols <= 65536; }

// Lays out H (CSR of the scaled matrix; H.perm maps to COO positions) on the
// device, as ELL when every row has the same length 1..16.
void build_dev_matrix(DevMatrix& D, const Csr& H, const double* unscaled_coo, bool want32,
                      bool want64, bool wantu, hipStream_t s) {
    D.rows = H.rows;
    D.cols = H.cols;
    D.nnz = (int)H.idx.size();
    int w = -1;
    bool uniform = H.rows > 0;
    for (int r = 0; r < H.rows && uniform; ++r) {
        const int len = H.ptr[r + 1] - H.ptr[r];
        if (w < 0) w = len;
        uniform = (len == w);
    }
    D.ell_w = (uniform && w >= 1 && w <= kMaxEllWidth) ? w : 0;
    const size_t nnz = H.idx.size();
    std::vector<int32_t> idx(nnz);
    std::vector<size_t> src(nnz);   // slot -> CSR position
    if (D.ell_w > 0) {
        for (int r = 0; r < H.rows; ++r)
            for (int k = 0; k < D.ell_w; ++k) src[(size_t)k * H.rows + r] = (size_t)H.ptr[r] + k;
    } else {
        for (size_t p = 0; p < nnz; ++p) src[p] = p;
        D.ptr = upload(H.ptr, s);
    }
    for (size_t q = 0; q < nnz; ++q) idx[q] = H.idx[src[q]];
    D.idx = upload(idx, s);
    if (narrow_ell(D)) {
        std::vector<uint16_t> i16(nnz);
        for (size_t q = 0; q < nnz; ++q) i16[q] = (uint16_t)idx[q];
        D.idx16 = upload(i16, s);
    }
    if (want32) {
        std::vector<float> v(nnz);
        for (size_t q = 0; q < nnz; ++q) v[q] = (float)H.val[src[q]];
        D.v32 = upload(v, s);
    }
    if (want64) {
        std::vector<double> v(nnz);
        for (size_t q = 0; q < nnz; ++q) v[q] = H.val[src[q]];
        D.v64 = upload(v, s);
    }
    if (wantu) {
        std::vector<double> v(nnz);
        for (size_t q = 0; q < nnz; ++q) v[q] = unscaled_coo[H.perm[src[q]]];
        D.vu = upload(v, s);
    }
    SUPR_HIP(hipStreamSynchronize(s));   // host staging vectors die here
}

// The same layout from a pattern built on the device (DevLp): scaled and
// unscaled values in CSR order; the pattern and value buffers are taken over
// (moved) where the CSR layout keeps them.
void build_dev_matrix(DevMatrix& D, DevPattern& P, DevBuf& scaled, DevBuf& unscaled, bool want32,
                      bool want64, bool wantu, hipStream_t s) {
    D.rows = P.rows;
    D.cols = P.cols;
    D.nnz = P.nnz;
    const bool uniform = P.rows > 0 && P.min_len == P.max_len;
    D.ell_w = (uniform && P.max_len >= 1 && P.max_len <= kMaxEllWidth) ? P.max_len : 0;
    const int64_t nnz = P.nnz;
    const int g = grid_elems(nnz);
    if (D.ell_w > 0) {
        auto ell = [&](auto* src, auto* dst) {
            hipLaunchKernelGGL(k_ell_from_csr, dim3(g), dim3(kBlock), 0, s, P.rows, D.ell_w, src, dst);
        };
        D.idx = device_vec<int32_t>(nnz);
        ell(P.idx.as<const int32_t>(), D.idx.as<int32_t>());
        if (narrow_ell(D)) {
            D.idx16 = device_vec<uint16_t>(nnz);
            ell(P.idx.as<const int32_t>(), D.idx16.as<uint16_t>());
        }
        if (want32) { D.v32 = device_vec<float>(nnz); ell(scaled.as<const double>(), D.v32.as<float>()); }
        if (want64) { D.v64 = device_vec<double>(nnz); ell(scaled.as<const double>(), D.v64.as<double>()); }
        if (wantu) { D.vu = device_vec<double>(nnz); ell(unscaled.as<const double>(), D.vu.as<double>()); }
        SUPR_HIP(hipGetLastError());
        SUPR_HIP(hipStreamSynchronize(s));
    } else {
        D.ptr = std::move(P.ptr);
        D.idx = std::move(P.idx);
        if (want32) {
            D.v32 = device_vec<float>(nnz);
            hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(kBlock), 0, s, nnz,
                               scaled.as<const double>(), D.v32.as<float>());
            SUPR_HIP(hipGetLastError());
            SUPR_HIP(hipStreamSynchronize(s));
        }
        if (want64) D.v64 = std::move(scaled);
        if (wantu) D.vu = std::move(unscaled);
    }
}

template <typename V>
void launch_spmv(const DevMatrix& M, const MatView<V>& view, const V* x, V* y, hipStream_t s) {
    SUPR_LANES(M.lanes, hipLaunchKernelGGL((k_spmv<LL, V, V, V>), dim3(grid_rows(M.rows, LL)),
                                           dim3(kBlock), 0, s, view, x, y));
    SUPR_HIP(hipGetLastError());
}

// Lanes per row for M from its shape alone (the default: deterministic, so
// every run of an LP sums in the same order and gives the same bits).  ELL
// rows take one lane, unless that leaves the matrix fewer than 65,536 lanes:
// then the smallest power of two that reaches 65,536, but no more than the
// power of two that covers the ELL width (no lane without a slot).  (A^T of
// the random 2,000 x 5,000 LP, 5,000 rows of 6: 8 lanes, 9.3 against
// 10.0 us per attempt at one lane.)  CSR rows take the larger of
//  * the largest power of two not above a quarter (FP32 values) or half
//    (FP64) of the mean row length (4 nonzeros or more per FP32 lane, 2 per
//    FP64 lane), and
//  * the smallest power of two that gives the matrix 65,536 lanes in all
//    (one 256-thread block per CU), so that a matrix with few rows still
//    fills the GPU,
// at most 32.  These are the timed autotuner's choices on the transportation
// (A: 2,000 rows of 1,000), random (A: 2,000 rows of 15) and power-law (A:
// 100,000-500,000 rows of 70; A^T: 2,000,000 rows of 17.6) LPs measured.
// Long rows that leave the GPU underfilled even at 32 lanes (fewer than 64
// blocks' worth, rows of 512 or more) take a block per row.
int rule_lanes(const DevMatrix& M, bool fp64) {
    if (M.rows <= 0) return 1;
    if (M.ell_w > 0) {
        int L = 1;
        while (L < M.ell_w && (double)M.rows * L < 65536.0) L *= 2;
        return L;
    }
    const double mean = (double)M.nnz / M.rows;
    const double t = fp64 ? mean / 2.0 : mean / 4.0;
    int L = 1;
    while (L < 32 && 2.0 * L <= t) L *= 2;
    while (L < 32 && (double)M.rows * L < 65536.0) L *= 2;
    if (L == 32 && mean >= 512.0 && (double)M.rows * L < 64.0 * kBlock) L = kBlock;
    return L;
}

// Picks lanes-per-row for M by timing every candidate on the device (the
// option value -1); the choice can change from run to run with the timing.
template <typename V>
int autotune_lanes(DevMatrix& M, const MatView<V>& view, const V* x, V* y, hipStream_t s) {
    if (M.ell_w > 0) return 1;
    const double mean = M.rows > 0 ? (double)M.nnz / M.rows : 0.0;
    hipEvent_t e0, e1;
    SUPR_HIP(hipEventCreate(&e0));
    SUPR_HIP(hipEventCreate(&e1));
    int best = 1;
    float best_ms = std::numeric_limits<float>::infinity();
    for (int L = 1; L <= kBlock; L = (L == 64 ? kBlock : 2 * L)) {
        if (L > 1 && mean < 0.25 * L) break;   // rows far shorter than the group
        M.lanes = L;
        launch_spmv(M, view, x, y, s);
        SUPR_HIP(hipEventRecord(e0, s));
        const int reps = 5;
        for (int r = 0; r < reps; ++r) launch_spmv(M, view, x, y, s);
        SUPR_HIP(hipEventRecord(e1, s));
        SUPR_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SUPR_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) { best_ms = ms; best = L; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    M.lanes = best;
    return best;
}

struct StreamGuard {
    hipStream_t s = nullptr;
    StreamGuard() { SUPR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~StreamGuard() { if (s) (void)hipStreamDestroy(s); }
};

// ---------------------------------------------------------------- solver --

double finite_inf_norm(const double* v, int len) {
    double m = 0.0;
    for (int k = 0; k < len; ++k)
        if (std::isfinite(v[k])) m = std::max(m, std::fabs(v[k]));
    return m;
}

double merit_of(const double* e) { return std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]); }

// Deep copy of a caller's LP, so an engine outlives the caller's arrays.
// Without the matrix, the COO pointers are null (the vectors are copied).
struct OwnedLp {
    std::vector<int32_t> row, col;
    std::vector<double> val, c, rL, rU, l, u;
    supr_lp v;
    struct Borrow {};
    // v aliases the caller's arrays (a one-shot solve: they outlive the solver)
    OwnedLp(const supr_lp& s, Borrow) : v(s) {}
    explicit OwnedLp(const supr_lp& s, bool with_matrix = true)
        : row(with_matrix ? s.row_idx : nullptr, with_matrix ? s.row_idx + s.nnz : nullptr),
          col(with_matrix ? s.col_idx : nullptr, with_matrix ? s.col_idx + s.nnz : nullptr),
          val(with_matrix ? s.val : nullptr, with_matrix ? s.val + s.nnz : nullptr),
          c(s.c, s.c + s.n), rL(s.rL, s.rL + s.m), rU(s.rU, s.rU + s.m), l(s.l, s.l + s.n),
          u(s.u, s.u + s.n), v(s) {
        v.row_idx = with_matrix ? row.data() : nullptr;
        v.col_idx = with_matrix ? col.data() : nullptr;
        v.val = with_matrix ? val.data() : nullptr;
        v.c = c.data(); v.rL = rL.data(); v.rU = rU.data(); v.l = l.data(); v.u = u.data();
    }
    OwnedLp(const OwnedLp&) = delete;
    OwnedLp& operator=(const OwnedLp&) = delete;
};

// Column relabelling for skewed column degrees (power-law LPs).  The dual
// half gathers x at the columns of every row of A; with skewed degrees a few
// hot columns take most of the gathers.  Sorting the columns by degree
// (descending, stable) packs the hot ones into few cache lines, and dealing
// the sorted columns out in groups of 16 (one 128-byte line of FP64 x)
// round-robin over the primal schedule's 256-column tiles keeps every tile a
// mix of degrees, so the primal's rows stay balanced (power-law LP
// 500,000 x 2,000,000: dual kernel 0.49 -> 0.30 ms, 804 -> 634 us per mixed
// attempt).  Applied when there are at least 2^16 columns and the largest
// degree is at least 4x the mean (SUPR_COLUMN_ORDER=given in the environment
// turns it off; the plain FP64 mode never relabels); the solver sees the
// relabelled forms and x is mapped back, so callers see the LP as given.
constexpr int32_t kRelabelMinCols = 1 << 16;
constexpr int kRelabelGroup = 16;

struct ColumnRelabel {
    std::vector<int32_t> newpos;   // old column -> new column
    std::vector<int32_t> col;      // relabelled COO column indices (the shared pattern)
    std::vector<double> uc, ul, uu, sc, sl, su, dc;
    supr_lp U{}, S{};              // views of the relabelled forms
};

bool relabel_columns(const supr_lp& U, const supr_lp& S, const double* Dc, int mode,
                     ColumnRelabel& R) {
    const int32_t n = S.n;
    const int64_t nnz = S.nnz;
    // the plain FP64 mode is the reference GPU engine's design, columns as given
    if (mode == SUPR_MODE_PLAIN_FP64) return false;
    if (n < kRelabelMinCols || nnz == 0) return false;
    if (const char* e = std::getenv("SUPR_COLUMN_ORDER"); e && std::strcmp(e, "given") == 0) return false;
    std::vector<int32_t> deg(n, 0);
    for (int64_t k = 0; k < nnz; ++k) ++deg[S.col_idx[k]];
    int32_t mx = 0;
    for (int32_t d : deg) mx = std::max(mx, d);
    if ((double)mx < 4.0 * (double)nnz / (double)n) return false;
    // stable counting sort by degree, descending (bucket mx - d)
    std::vector<int64_t> start((size_t)mx + 2, 0);
    for (int32_t d : deg) ++start[(size_t)(mx - d) + 1];
    for (size_t b = 1; b < start.size(); ++b) start[b] += start[b - 1];
    std::vector<int32_t> srt(n);
    for (int32_t j = 0; j < n; ++j) srt[start[(size_t)(mx - deg[j])]++] = j;
    // group g of the whole tiles goes to tile g % T, slot g / T; the rest
    // (under one tile) keeps its sorted place at the end
    constexpr int tile = kBlock;
    const int64_t T = (int64_t)n / tile;
    const int64_t ng = T * (tile / kRelabelGroup);
    R.newpos.resize(n);
    for (int64_t g = 0; g < ng; ++g) {
        const int64_t base = (g % T) * tile + (g / T) * kRelabelGroup;
        for (int k = 0; k < kRelabelGroup; ++k) R.newpos[srt[g * kRelabelGroup + k]] = (int32_t)(base + k);
    }
    for (int64_t i = ng * kRelabelGroup; i < n; ++i) R.newpos[srt[i]] = (int32_t)i;
    R.col.resize(nnz);
    for (int64_t k = 0; k < nnz; ++k) R.col[k] = R.newpos[S.col_idx[k]];
    auto perm = [&](const double* v, std::vector<double>& out) {
        out.resize(n);
        for (int32_t j = 0; j < n; ++j) out[R.newpos[j]] = v[j];
    };
    perm(U.c, R.uc); perm(U.l, R.ul); perm(U.u, R.uu);
    perm(S.c, R.sc); perm(S.l, R.sl); perm(S.u, R.su);
    perm(Dc, R.dc);
    R.U = U;
    R.U.col_idx = R.col.data(); R.U.c = R.uc.data(); R.U.l = R.ul.data(); R.U.u = R.uu.data();
    R.S = S;
    R.S.col_idx = R.col.data(); R.S.c = R.sc.data(); R.S.l = R.sl.data(); R.S.u = R.su.data();
    return true;
}

// x in the caller's column order from the relabelled solve's xp (the
// objective is the solver's: the same terms, summed by the device's check).
void map_back(const std::vector<int32_t>& newpos, const double* xp, double* x_out,
              const supr_pdlp_info* info) {
    if (!info->has_point) return;   // x_out untouched, as the solver leaves it
    const size_t n = newpos.size();
    for (size_t j = 0; j < n; ++j) x_out[j] = xp[newpos[j]];
}

class Solver {
   public:
    Solver(const supr_lp& U, const supr_lp& S, const double* Dr, const double* Dc,
           const supr_pdlp_options& o)
        : Uown_(U), Sown_(S), Drv_(Dr, Dr + S.m), Dcv_(Dc, Dc + S.n), U_(Uown_.v), S_(Sown_.v),
          Dr_(Drv_.data()), Dc_(Dcv_.data()), o_(o), m_(S.m), n_(S.n) {}
    // A one-shot solve: the caller's forms and factors are read in place, not
    // copied (they are const and outlive the call).
    Solver(const supr_lp& U, const supr_lp& S, const double* Dr, const double* Dc,
           const supr_pdlp_options& o, OwnedLp::Borrow b)
        : Uown_(U, b), Sown_(S, b), U_(Uown_.v), S_(Sown_.v), Dr_(Dr), Dc_(Dc), o_(o), m_(S.m),
          n_(S.n) {}
    // From the unscaled form alone: the matrix is laid out and scaled on the
    // device (scale(), src/scaling.cpp:25-84, bit-identical to the host's);
    // the scaled vectors and Dr, Dc come back to the host (n + m values).
    Solver(const supr_lp& U, const char* integer, int ruiz_passes, const supr_pdlp_options& o)
        : Uown_(U, o.mode == SUPR_MODE_PLAIN_FP64), Sown_(U, false), Drv_(U.m), Dcv_(U.n),
          U_(Uown_.v), S_(Sown_.v), Dr_(Drv_.data()), Dc_(Dcv_.data()), o_(o), m_(U.m), n_(U.n),
          device_prep_(true), passes_(ruiz_passes) {
        if (integer) integer_.assign(integer, integer + U.n);
        dlp_ = std::make_unique<DevLp>(U, stream_.s);   // uploads the caller's arrays now
    }
    // From a form already on the device (built there from the original LP);
    // Uvec carries the form's sizes and host vectors (null COO pointers).
    Solver(std::unique_ptr<DevLp> form, const supr_lp& Uvec, const supr_pdlp_options& o)
        : Uown_(Uvec, false), Sown_(Uvec, false), Drv_(Uvec.m), Dcv_(Uvec.n), U_(Uown_.v),
          S_(Sown_.v), Dr_(Drv_.data()), Dc_(Dcv_.data()), o_(o), m_(Uvec.m), n_(Uvec.n),
          device_prep_(true), passes_(10) {
        if (o.mode == SUPR_MODE_PLAIN_FP64)
            throw BadInput("the plain FP64 mode takes host forms (supr_solve_pdhg_gpu)");
        dlp_ = std::move(form);
    }
    void set_tol(double tol) { o_.tol = tol; }
    ~Solver() {
        if (stream_.s) (void)hipStreamSynchronize(stream_.s);
        for (auto& g : graph_) if (g) (void)hipGraphExecDestroy(g);
        if (mirror_) (void)hipHostFree(mirror_);
    }

    // Uploads the LP, lays out the matrices, tunes the SpMV schedules.
    void setup();
    // One solve from the cold start x = clip(0, l, u), y = 0 (pdlp_algo.hpp:67-69).
    void run(supr_stop_fn stop, void* ctx, double* x_out, double* y_out, supr_pdlp_info* info);
    void bench_kernels(int reps, double* out);
    double setup_seconds() const { return setup_s_; }

   private:
    OwnedLp Uown_, Sown_;
    std::vector<double> Drv_, Dcv_;
    const supr_lp& U_;
    const supr_lp& S_;
    const double* Dr_;
    const double* Dc_;
    supr_pdlp_options o_;
    const int m_, n_;
    StreamGuard stream_;
    DevMatrix A_, AT_;
    int lanesA_[2] = {1, 1}, lanesAT_[2] = {1, 1};   // [0] FP32, [1] FP64 schedules
    double eta0_ = 1.0, omega0_ = 1.0;
    std::vector<double> x0_;
    double setup_s_ = 0.0;
    bool ready_ = false;

    // device preprocessing (the unscaled-form constructor)
    bool device_prep_ = false;
    // one-pass iteration kernels where the launch covers every row in one
    // pass (SUPR_ONE_PASS=off: the grid-stride kernels everywhere; same bits)
    bool one_pass_ = true;
    // ... and where the prefetched group holds a mean row (kPreU slots per
    // lane): on long rows the prefetch only adds registers (the 1M-variable
    // LP's A, 1,000 per row at 32 lanes: 61.4 against 59.6 ms per solve)
    bool one_pass(const DevMatrix& M) const {
        return one_pass_ && (long)M.nnz <= (long)kPreU * M.lanes * M.rows;
    }
    // The dual launches' form (launch_dual): the one-pass form where one
    // prefetched group holds a mean row (1); for long CSR rows the flat
    // one-pass form (2: each lane's entries in chunks of 32, the first loaded
    // before the state, every chunk's gathers issued together), on grids of
    // at most 1,024 blocks (four wavefronts per SIMD), where its registers
    // cost no residency; otherwise the grid-stride kernel (0).
    int one_pass_dual(const DevMatrix& M) const {
        if (one_pass(M)) return 1;
        return one_pass_ && M.ell_w == 0 && grid_rows(M.rows, M.lanes) <= 1024 ? 2 : 0;
    }
    // two adjacent rows of A^T per lane in the primal launches where the
    // layout allows it (primal_view; SUPR_ROW_PAIRS=off: one row per lane)
    bool row_pairs_ = true;
    // FP64 row pairs in one pass take k_primal_pairs (SUPR_EARLY_SLOTS=off:
    // k_primal's row-pair path; the same bits)
    bool early_slots_ = true;
    int passes_ = 10;
    std::vector<char> integer_;
    std::unique_ptr<DevLp> dlp_;
    std::unique_ptr<DevScaled> dsc_;
    double amax_ = 0.0;                 // max |scaled a_ij|, from the device
    std::unique_ptr<OwnedLp> uhost_;    // U with its matrix, fetched when needed
    void prepare_on_device();
    const supr_lp& unscaled_with_matrix();

    void initial_point();
    void setup_matrices(bool want32, bool want64, bool wantu);
    void use_lanes(int prec) { A_.lanes = lanesA_[prec]; AT_.lanes = lanesAT_[prec]; }

    // plain mode
    DevBuf pc_, pl_, pu_, pb_, pineq_, px_, pxn_, paty_, pxsum_, py_, pyn_, pax_, paxn_, pysum_, pstats_;
    std::unique_ptr<Csr> Au_, AuT_;
    void setup_plain();
    void run_plain(supr_stop_fn stop, void* ctx, double* x_out, double* y_out, supr_pdlp_info* info);

    // device-resident modes
    template <typename T> struct Iter {
        DevBuf x[2], y[2], ax[2], xsum, ysum, xlast, ylast, xd, c, l, u, b;
        // every l_j is +0 and every u_j is +inf in this precision: l and u are
        // not stored and the kernels project onto x >= 0
        bool nonneg = false;
        IterBufs<T> view(const uint8_t* ineq) const {
            IterBufs<T> B;
            for (int k = 0; k < 2; ++k) {
                B.x[k] = x[k].template as<T>();
                B.y[k] = y[k].template as<T>();
                B.ax[k] = ax[k].template as<T>();
            }
            B.aty = nullptr;   // formed by every primal launch, never stored
            B.xsum = xsum.template as<T>();
            B.ysum = ysum.template as<T>();
            B.xlast = xlast.template as<T>();
            B.ylast = ylast.template as<T>();
            B.xd = xd.template as<T>();
            B.c = c.template as<T>();
            B.l = nonneg ? nullptr : l.template as<T>();
            B.u = nonneg ? nullptr : u.template as<T>();
            B.b = b.template as<T>();
            B.ineq = ineq;
            return B;
        }
    };
    Iter<float> F_;
    Iter<double> D_;
    // one allocation holds every iterate, KKT and scratch buffer below (views)
    Arena arena_;
    std::vector<uint8_t> hineq_;   // inequality flags, host copy of ineq_
    DevBuf ineq_, state_, parts_, x0d_;
    DevBuf ellc_;   // (Dr_i a) per row of A, when A^T's FP64 values are formed from factors
    void rebuild_ell_values();
    DevBuf kc_, kl_, ku_, krL_, krU_, kDr_, kDc_, xu_, xu2_, xavg_, yu_, yu2_, yavg_, bestx_, besty_,
        axavg_;
    KktBufs K_{};
    double *part_n_ = nullptr, *part_m_ = nullptr, *part_kn_ = nullptr, *part_km_ = nullptr,
           *part_rn_ = nullptr, *part_rm_ = nullptr, *dx2_slot_ = nullptr;
    hipGraphExec_t graph_[2] = {nullptr, nullptr};   // per iterate precision
    DevState* mirror_ = nullptr;       // host-coherent state copy (host pointer)
    DevState* mirror_dev_ = nullptr;   // the same memory, device address

    template <typename T> void take_iter(Arena& a, Iter<T>& I);
    void take_buffers(Arena& a, bool mixed);
    template <typename T> void reset_iter(Iter<T>& I, const DevState* st = nullptr);
    void setup_device_resident(bool mixed);
    DevState initial_state(bool mixed) const;
    template <typename T> void enqueue_attempts(const Iter<T>& I, int count);
    template <typename T> void enqueue_check(const Iter<T>& I);
    template <typename T> void set_products(const Iter<T>& I, const MatView<T>& A);
    template <typename T> hipGraphExec_t capture_graph(const Iter<T>& I, int attempts, int units);
    template <typename T>
    int run_graph_loop(const Iter<T>& I, int prec, supr_stop_fn stop, void* ctx, DevState& hs,
                       bool& stopped);
    void run_device_resident(bool mixed, supr_stop_fn stop, void* ctx, double* x_out,
                             double* y_out, supr_pdlp_info* info);
};

void Solver::initial_point() {
    // include/pdlp_algo.hpp:67-77: x0 = clip(0, l, u), eta0 = 1 / max|a_ij|,
    // omega0 = |c| / |finite rL| (1 when either is tiny).
    x0_.resize(n_);
    for (int j = 0; j < n_; ++j) x0_[j] = smin(smax(0.0, S_.l[j]), S_.u[j]);
    double amax = amax_;
    if (!device_prep_)
        for (int k = 0; k < S_.nnz; ++k) amax = smax(amax, std::fabs(S_.val[k]));
    eta0_ = amax > 0 ? 1.0 / amax : 1.0;
    double cn = 0.0, bn = 0.0;
    for (int j = 0; j < n_; ++j) cn += S_.c[j] * S_.c[j];
    for (int i = 0; i < m_; ++i)
        if (std::isfinite(S_.rL[i])) bn += S_.rL[i] * S_.rL[i];
    cn = std::sqrt(cn);
    bn = std::sqrt(bn);
    omega0_ = (cn > 1e-10 && bn > 1e-10) ? cn / bn : 1.0;
}

// Caller-scaled forms of at least this many nonzeros are laid out on the
// device (1M-variable transportation LP: setup 65 -> 22 ms), unless the
// environment sets SUPR_HOST_LAYOUT=1 (the host build, for comparison).
constexpr int32_t kDeviceLayoutNnz = 1 << 16;
bool host_layout_forced() {
    const char* e = std::getenv("SUPR_HOST_LAYOUT");
    return e && e[0] == '1';
}

void Solver::setup_matrices(bool want32, bool want64, bool wantu) {
    if (device_prep_) {
        build_dev_matrix(A_, dlp_->A, dsc_->vA, dlp_->vA, want32, want64, wantu, stream_.s);
        build_dev_matrix(AT_, dlp_->AT, dsc_->vAT, dlp_->vAT, want32, want64, wantu, stream_.s);
        // keep only the COO arrays (the host verifier's fallback input)
        dsc_.reset();
        dlp_->A = DevPattern();
        dlp_->AT = DevPattern();
        dlp_->vA.reset();
        dlp_->vAT.reset();
        return;
    }
    if (o_.mode != SUPR_MODE_PLAIN_FP64 && S_.nnz >= kDeviceLayoutNnz && !host_layout_forced()) {
        // The caller's two forms share one pattern: upload it once and lay out
        // A and A^T on the device (the same stable counting sort as the host's
        // build_csr, so the same layout bit for bit), with the scaled values
        // and the unscaled ones gathered into both orders.  Below
        // kDeviceLayoutNnz nonzeros the host build is quicker than the launches.
        hipStream_t s = stream_.s;
        const int32_t nnz = S_.nnz;
        DevLp dl(m_, n_, nnz, upload(S_.row_idx, nnz, s), upload(S_.col_idx, nnz, s),
                 upload(U_.val, nnz, s), DevBuf(), DevBuf(), DevBuf(), DevBuf(), DevBuf(), s);
        const DevBuf sv = upload(S_.val, nnz, s);
        DevBuf svA, svAT;
        gather_values(dl.A, sv.as<double>(), svA, s);
        gather_values(dl.AT, sv.as<double>(), svAT, s);
        build_dev_matrix(A_, dl.A, svA, dl.vA, want32, want64, wantu, s);
        build_dev_matrix(AT_, dl.AT, svAT, dl.vAT, want32, want64, wantu, s);
        SUPR_HIP(hipStreamSynchronize(s));   // dl and sv are freed on return
        return;
    }
    const Csr A = build_csr(S_, S_.val, false);
    build_dev_matrix(A_, A, U_.val, want32, want64, wantu, stream_.s);
    const Csr AT = build_csr(S_, S_.val, true);
    build_dev_matrix(AT_, AT, U_.val, want32, want64, wantu, stream_.s);
}

// scale() on the device; the scaled vectors, Dr and Dc come back to the host
// (they seed the initial point and the unscaling of the results).
void Solver::prepare_on_device() {
    hipStream_t s = stream_.s;
    dsc_ = std::make_unique<DevScaled>();
    scale_on_device(*dlp_, integer_.empty() ? nullptr : integer_.data(), passes_, *dsc_, s);
    amax_ = device_absmax(dsc_->vA.as<double>(), dlp_->nnz, s);
    auto down = [&](double* dst, const DevBuf& src, int len) {
        if (len > 0)
            SUPR_HIP(hipMemcpyAsync(dst, src.as<void>(), sizeof(double) * len, hipMemcpyDeviceToHost, s));
    };
    down(Sown_.c.data(), dsc_->c, n_); down(Sown_.l.data(), dsc_->l, n_); down(Sown_.u.data(), dsc_->u, n_);
    down(Sown_.rL.data(), dsc_->rL, m_); down(Sown_.rU.data(), dsc_->rU, m_);
    down(Drv_.data(), dsc_->Dr, m_); down(Dcv_.data(), dsc_->Dc, n_);
    SUPR_HIP(hipStreamSynchronize(s));
}

// The unscaled form with its COO matrix on the host (the caller's copy, or
// fetched from the device for an engine built from the unscaled form).
const supr_lp& Solver::unscaled_with_matrix() {
    if (U_.nnz == 0 || U_.val != nullptr) return U_;
    if (!uhost_) {
        supr_lp shape = U_;
        std::vector<int32_t> r(U_.nnz), c(U_.nnz);
        std::vector<double> v(U_.nnz);
        hipStream_t s = stream_.s;
        SUPR_HIP(hipMemcpyAsync(r.data(), dlp_->row.as<void>(), sizeof(int32_t) * U_.nnz, hipMemcpyDeviceToHost, s));
        SUPR_HIP(hipMemcpyAsync(c.data(), dlp_->col.as<void>(), sizeof(int32_t) * U_.nnz, hipMemcpyDeviceToHost, s));
        SUPR_HIP(hipMemcpyAsync(v.data(), dlp_->val.as<void>(), sizeof(double) * U_.nnz, hipMemcpyDeviceToHost, s));
        SUPR_HIP(hipStreamSynchronize(s));
        shape.row_idx = r.data(); shape.col_idx = c.data(); shape.val = v.data();
        uhost_ = std::make_unique<OwnedLp>(shape);
    }
    return uhost_->v;
}

void Solver::setup() {
    const auto t0 = Clock::now();
    if (device_prep_) prepare_on_device();
    switch (o_.mode) {
        case SUPR_MODE_PLAIN_FP64: setup_plain(); break;
        case SUPR_MODE_FAST_FP64: setup_device_resident(false); break;
        case SUPR_MODE_MIXED: setup_device_resident(true); break;
        default: throw BadInput("unknown mode " + std::to_string(o_.mode));
    }
    ready_ = true;
    setup_s_ = seconds_since(t0);
}

void Solver::run(supr_stop_fn stop, void* ctx, double* x_out, double* y_out,
                 supr_pdlp_info* info) {
    if (!ready_) throw BadInput("engine is not set up");
    const auto t0 = Clock::now();
    std::memset(info, 0, sizeof(*info));
    if (o_.mode == SUPR_MODE_PLAIN_FP64) run_plain(stop, ctx, x_out, y_out, info);
    else run_device_resident(o_.mode == SUPR_MODE_MIXED, stop, ctx, x_out, y_out, info);
    info->seconds_setup = setup_s_;
    info->seconds_total = seconds_since(t0) + setup_s_;
}

// ---------------------------------------------------------- plain FP64 --

void Solver::setup_plain() {
    hipStream_t s = stream_.s;
    const int n = n_, m = m_;
    setup_matrices(false, true, false);
    initial_point();
    // host CSR of the unscaled form for the host verifier (pdlp_algo.hpp:57)
    Au_ = std::make_unique<Csr>(build_csr(U_, U_.val, false));
    AuT_ = std::make_unique<Csr>(build_csr(U_, U_.val, true));
    auto vec = [&](const double* p, int len) { return upload(p, (size_t)len, s); };
    pc_ = vec(S_.c, n); pl_ = vec(S_.l, n); pu_ = vec(S_.u, n); pb_ = vec(S_.rL, m);
    std::vector<uint8_t> hineq(m);
    for (int i = 0; i < m; ++i) hineq[i] = S_.rL[i] != S_.rU[i];
    pineq_ = upload(hineq, s);
    px_ = device_vec<double>(n); pxn_ = device_vec<double>(n); paty_ = device_vec<double>(n);
    pxsum_ = device_vec<double>(n); py_ = device_vec<double>(m); pyn_ = device_vec<double>(m);
    pax_ = device_vec<double>(m); paxn_ = device_vec<double>(m); pysum_ = device_vec<double>(m);
    pstats_ = device_vec<double>(3 * kMaxPlain);
    if (o_.spmv_lanes_A > 0) A_.lanes = o_.spmv_lanes_A;
    else if (o_.spmv_lanes_A == 0) A_.lanes = rule_lanes(A_, true);
    else autotune_lanes(A_, A_.f64(), px_.as<double>(), pax_.as<double>(), s);
    if (o_.spmv_lanes_AT > 0) AT_.lanes = o_.spmv_lanes_AT;
    else if (o_.spmv_lanes_AT == 0) AT_.lanes = rule_lanes(AT_, true);
    else autotune_lanes(AT_, AT_.f64(), py_.as<double>(), paty_.as<double>(), s);
    if (A_.ell_w > 0 && o_.spmv_lanes_A <= 0) A_.lanes = 1;
    if (AT_.ell_w > 0 && o_.spmv_lanes_AT <= 0) AT_.lanes = 1;
    lanesA_[1] = A_.lanes; lanesAT_[1] = AT_.lanes;
}

// The reference's GpuBackend loop (include/pdlp_algo.hpp:103-175 with
// gpu/pdlp_gpu_solve.cu): step statistics read back after every attempt,
// KKT check on the host.
void Solver::run_plain(supr_stop_fn stop, void* ctx, double* x_out, double* y_out,
                       supr_pdlp_info* info) {
    hipStream_t s = stream_.s;
    const int n = n_, m = m_;
    use_lanes(1);
    double* px = px_.as<double>(); double* pxn = pxn_.as<double>();
    double* py = py_.as<double>(); double* pyn = pyn_.as<double>();
    double* pax = pax_.as<double>(); double* paxn = paxn_.as<double>();
    double* aty = paty_.as<double>(); double* xsum = pxsum_.as<double>(); double* ysum = pysum_.as<double>();
    std::vector<double> hx = x0_, hy(m, 0.0);
    SUPR_HIP(hipMemcpyAsync(px, hx.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    SUPR_HIP(hipMemsetAsync(py, 0, sizeof(double) * m, s));
    SUPR_HIP(hipMemsetAsync(xsum, 0, sizeof(double) * n, s));
    SUPR_HIP(hipMemsetAsync(ysum, 0, sizeof(double) * m, s));
    launch_spmv(A_, A_.f64(), px, pax, s);
    launch_spmv(AT_, AT_.f64(), py, aty, s);
    SUPR_HIP(hipStreamSynchronize(s));
    const auto t1 = Clock::now();

    double eta = eta0_, omega = omega0_, wsum = 0.0;
    std::vector<double> hx_last = hx, hy_last = hy, hxs(n), hys(m), xu(n), yu(m), xu2(n), yu2(m);
    double merit_last_restart = INFINITY, merit_prev_cand = INFINITY;
    int since_restart = 0, restarts = 0, accepted = 0, attempts = 0, checks = 0;
    std::vector<double> best_x, best_y;
    double best_merit = INFINITY, best_eps[3] = {0, 0, 0};
    double stats_h[3 * kMaxPlain];
    const int gb_n = (n + kBlock - 1) / kBlock, gb_m = (m + kBlock - 1) / kBlock;
    const int rb_n = std::min(kMaxPlain, std::max(1, gb_n)), rb_m = std::min(kMaxPlain, std::max(1, gb_m));
    double rc[7], ra[7];
    bool converged = false;
    std::vector<double> res_x, res_y;
    double res_eps[3] = {0, 0, 0};
    const double* c = pc_.as<double>(); const double* l = pl_.as<double>();
    const double* u = pu_.as<double>(); const double* b = pb_.as<double>();
    const uint8_t* ineq = pineq_.as<uint8_t>();

    auto unscaled_kkt = [&](const std::vector<double>& xs, const std::vector<double>& ys,
                            std::vector<double>& xuo, std::vector<double>& yuo, double* rep) {
        for (int j = 0; j < n; ++j) xuo[j] = Dc_[j] * xs[j];
        for (int i = 0; i < m; ++i) yuo[i] = Dr_[i] * ys[i];
        verify_csr(U_, *Au_, *AuT_, xuo.data(), yuo.data(), rep);
    };
    auto keep_best = [&](const std::vector<double>& xv, const std::vector<double>& yv, const double* r) {
        const double mm = merit_of(r);
        if (mm < best_merit) {
            best_merit = mm; best_x = xv; best_y = yv;
            best_eps[0] = r[0]; best_eps[1] = r[1]; best_eps[2] = r[2];
        }
    };

    while (attempts < o_.max_iterations) {
        if (stop && stop(ctx)) break;
        while (attempts < o_.max_iterations) {
            ++attempts;
            const double tau = eta / omega, sigma = eta * omega;
            hipLaunchKernelGGL(k_plain_primal, dim3(std::max(1, gb_n)), dim3(kBlock), 0, s, n, px, c, aty, l, u, tau, pxn);
            launch_spmv(A_, A_.f64(), pxn, paxn, s);
            hipLaunchKernelGGL(k_plain_dual, dim3(std::max(1, gb_m)), dim3(kBlock), 0, s, m, py, b, paxn, pax, ineq, sigma, pyn);
            hipLaunchKernelGGL(k_plain_stats_n, dim3(rb_n), dim3(kBlock), 0, s, n, px, pxn, pstats_.as<double>());
            hipLaunchKernelGGL(k_plain_stats_m, dim3(rb_m), dim3(kBlock), 0, s, m, py, pyn, pax, paxn, pstats_.as<double>());
            SUPR_HIP(hipGetLastError());
            SUPR_HIP(hipMemcpyAsync(stats_h, pstats_.as<double>(), sizeof(stats_h), hipMemcpyDeviceToHost, s));
            SUPR_HIP(hipStreamSynchronize(s));
            double dx2 = 0.0, dy2 = 0.0, inter = 0.0;
            for (int k = 0; k < rb_n; ++k) dx2 += stats_h[k];
            for (int k = 0; k < rb_m; ++k) { dy2 += stats_h[kMaxPlain + k]; inter += stats_h[2 * kMaxPlain + k]; }
            const double num = omega * dx2 + dy2 / omega;
            const double eta_bar = std::fabs(inter) > 0 ? num / (2.0 * std::fabs(inter)) : INFINITY;
            const double k1 = accepted + 2.0;
            const double grow = (1.0 + std::pow(k1, -0.6)) * eta;
            const double eta_next = std::isfinite(eta_bar) ? smin((1.0 - std::pow(k1, -0.3)) * eta_bar, grow) : grow;
            if (eta <= eta_bar) {
                std::swap(px, pxn); std::swap(py, pyn); std::swap(pax, paxn);
                launch_spmv(AT_, AT_.f64(), py, aty, s);
                hipLaunchKernelGGL(k_plain_axpy, dim3(std::max(1, gb_n)), dim3(kBlock), 0, s, n, eta, px, xsum);
                hipLaunchKernelGGL(k_plain_axpy, dim3(std::max(1, gb_m)), dim3(kBlock), 0, s, m, eta, py, ysum);
                wsum += eta;
                eta = eta_next;
                ++accepted; ++since_restart;
                break;
            }
            eta = eta_next;
        }
        if (accepted % o_.check_every != 0 && attempts < o_.max_iterations) continue;

        ++checks;
        SUPR_HIP(hipMemcpyAsync(hx.data(), px, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        SUPR_HIP(hipMemcpyAsync(hy.data(), py, sizeof(double) * m, hipMemcpyDeviceToHost, s));
        SUPR_HIP(hipMemcpyAsync(hxs.data(), xsum, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        SUPR_HIP(hipMemcpyAsync(hys.data(), ysum, sizeof(double) * m, hipMemcpyDeviceToHost, s));
        SUPR_HIP(hipStreamSynchronize(s));
        if (wsum > 0) { for (auto& v : hxs) v /= wsum; for (auto& v : hys) v /= wsum; }
        else { hxs = hx; hys = hy; }
        unscaled_kkt(hx, hy, xu, yu, rc);
        unscaled_kkt(hxs, hys, xu2, yu2, ra);
        keep_best(xu, yu, rc);
        keep_best(xu2, yu2, ra);
        const bool use_avg = merit_of(ra) < merit_of(rc);
        const double* rep = use_avg ? ra : rc;
        const double mc = merit_of(rep);
        if (smax(smax(rep[0], rep[1]), rep[2]) <= o_.tol) {
            res_x = use_avg ? xu2 : xu; res_y = use_avg ? yu2 : yu;
            res_eps[0] = rep[0]; res_eps[1] = rep[1]; res_eps[2] = rep[2];
            converged = true;
            break;
        }
        const bool sufficient = mc <= 0.2 * merit_last_restart;
        const bool necessary = mc <= 0.8 * merit_last_restart && mc > merit_prev_cand;
        const bool artificial = since_restart >= 0.36 * accepted;
        merit_prev_cand = mc;
        if (sufficient || necessary || artificial) {
            const std::vector<double>& cx = use_avg ? hxs : hx;
            const std::vector<double>& cy = use_avg ? hys : hy;
            double ddx = 0.0, ddy = 0.0;
            for (int j = 0; j < n; ++j) ddx += (cx[j] - hx_last[j]) * (cx[j] - hx_last[j]);
            for (int i = 0; i < m; ++i) ddy += (cy[i] - hy_last[i]) * (cy[i] - hy_last[i]);
            ddx = std::sqrt(ddx); ddy = std::sqrt(ddy);
            if (ddx > 1e-10 && ddy > 1e-10) omega = std::exp(0.5 * std::log(ddy / ddx) + 0.5 * std::log(omega));
            hx_last = cx; hy_last = cy;
            if (use_avg) {
                SUPR_HIP(hipMemcpyAsync(px, cx.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
                SUPR_HIP(hipMemcpyAsync(py, cy.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
                launch_spmv(A_, A_.f64(), px, pax, s);
                launch_spmv(AT_, AT_.f64(), py, aty, s);
            }
            SUPR_HIP(hipMemsetAsync(xsum, 0, sizeof(double) * n, s));
            SUPR_HIP(hipMemsetAsync(ysum, 0, sizeof(double) * m, s));
            SUPR_HIP(hipStreamSynchronize(s));   // cx / cy are host vectors reused below
            wsum = 0.0;
            since_restart = 0; ++restarts;
            merit_last_restart = mc;
            merit_prev_cand = INFINITY;
        }
    }
    SUPR_HIP(hipStreamSynchronize(s));
    bool has_point = converged;
    if (!converged) {
        if (!std::isfinite(best_merit)) {
            SUPR_HIP(hipMemcpy(hx.data(), px, sizeof(double) * n, hipMemcpyDeviceToHost));
            SUPR_HIP(hipMemcpy(hy.data(), py, sizeof(double) * m, hipMemcpyDeviceToHost));
            unscaled_kkt(hx, hy, xu, yu, rc);
            keep_best(xu, yu, rc);
        }
        has_point = std::isfinite(best_merit);
        res_x = best_x; res_y = best_y;
        std::copy(best_eps, best_eps + 3, res_eps);
    }
    if (has_point) {
        std::copy(res_x.begin(), res_x.end(), x_out);
        std::copy(res_y.begin(), res_y.end(), y_out);
    }
    info->iterations = attempts;
    info->restarts = restarts;
    info->converged = converged ? 1 : 0;
    info->has_point = has_point ? 1 : 0;
    info->eps_P = res_eps[0]; info->eps_D = res_eps[1]; info->eps_G = res_eps[2];
    double obj = U_.obj_offset;
    if (has_point) { obj = 0.0; for (int j = 0; j < n; ++j) obj += U_.c[j] * x_out[j]; obj += U_.obj_offset; }
    info->objective = obj;
    info->iterations_fp32 = 0;
    info->checks = checks;
    info->lanes_A = A_.lanes; info->lanes_AT = AT_.lanes;
    info->seconds_solve = seconds_since(t1);
}

// ---------------------------------------------------- device-resident ---

// Every l_j is +0 and every u_j is +inf in precision T (then the kernels
// project onto x >= 0 and read neither).
// Chunks of 4,096 without branches (vectorised), stopping after the first
// chunk that fails.
template <typename T> bool nonneg_bounds(const double* l, const double* u, int n) {
    constexpr int C = 4096;
    for (int j0 = 0; j0 < n; j0 += C) {
        const int j1 = std::min(n, j0 + C);
        bool ok = true;
        for (int j = j0; j < j1; ++j) {
            const T lj = (T)l[j], uj = (T)u[j];
            ok &= (lj == T(0)) & !std::signbit(lj) & (uj == std::numeric_limits<T>::infinity());
        }
        if (!ok) return false;
    }
    return true;
}

// The iterate buffers of one precision, as views into the arena.
template <typename T>
void Solver::take_iter(Arena& a, Iter<T>& I) {
    const size_t n = n_, m = m_;
    for (int k = 0; k < 2; ++k) {
        I.x[k] = a.take<T>(n);
        I.y[k] = a.take<T>(m);
        I.ax[k] = a.take<T>(m);
    }
    I.xsum = a.take<T>(n);
    I.ysum = a.take<T>(m);
    I.xlast = a.take<T>(n);
    I.ylast = a.take<T>(m);
    if (sizeof(T) == 4) I.xd = a.take<T>(2 * n);   // (xn, xn - x) pairs
    I.c = a.take<T>(n);
    I.b = a.take<T>(m);
    if (!I.nonneg) {
        I.l = a.take<T>(n);
        I.u = a.take<T>(n);
    }
}

// Every buffer of the device-resident modes, in one fixed order (called once
// to count the bytes and once, after the arena's allocation, to place them).
void Solver::take_buffers(Arena& a, bool mixed) {
    const size_t n = n_, m = m_;
    ineq_ = a.take<uint8_t>(m);
    x0d_ = a.take<double>(n);
    if (mixed) take_iter(a, F_);
    take_iter(a, D_);
    kc_ = a.take<double>(n); kl_ = a.take<double>(n); ku_ = a.take<double>(n);
    krL_ = a.take<double>(m); krU_ = a.take<double>(m);
    kDr_ = a.take<double>(m); kDc_ = a.take<double>(n);
    xu_ = a.take<double>(n); xu2_ = a.take<double>(n); xavg_ = a.take<double>(n);
    yu_ = a.take<double>(m); yu2_ = a.take<double>(m); yavg_ = a.take<double>(m);
    bestx_ = a.take<double>(n); besty_ = a.take<double>(m);
    axavg_ = a.take<double>(m);
    // partial-sum scratch: 1 + 2 + 8 + 6 + 1 + 1 slots per block, plus dx2
    parts_ = a.take<double>((size_t)19 * kMaxBlocks + 64);
    state_ = a.take<DevState>(2);   // double-buffered (see pdlp_kernels.hpp)
}

// Cold start: x = x0, y = 0, sums cleared, restart point = start point, and
// (st given) both halves of the state set: one launch on the stream, from
// the device copy of x0, with the state passed by value.
template <typename T>
void Solver::reset_iter(Iter<T>& I, const DevState* st) {
    const IterBufs<T> B = I.view(ineq_.as<uint8_t>());
    hipLaunchKernelGGL((k_reset<T>), dim3(grid_elems(std::max(n_, m_))), dim3(kBlock), 0, stream_.s, n_, m_,
                       x0d_.as<const double>(), B, st ? state_.as<DevState>() : nullptr,
                       st ? *st : DevState{});
    SUPR_HIP(hipGetLastError());
}

void Solver::setup_device_resident(bool mixed) {
    hipStream_t s = stream_.s;
    const int n = n_, m = m_;
    const auto t0 = Clock::now();
    if (const char* e = std::getenv("SUPR_ONE_PASS"); e && std::strcmp(e, "off") == 0) one_pass_ = false;
    if (const char* e = std::getenv("SUPR_ROW_PAIRS"); e && std::strcmp(e, "off") == 0) row_pairs_ = false;
    if (const char* e = std::getenv("SUPR_EARLY_SLOTS"); e && std::strcmp(e, "off") == 0) early_slots_ = false;
    // The host passes over the vectors (x0, eta0, omega0; the row kinds and
    // the bound kinds) run on two worker threads while this thread uploads
    // and lays out the matrices; they only read the forms.
    bool unonneg = false;
    auto init = std::async(std::launch::async, [this] { initial_point(); });
    auto kinds = std::async(std::launch::async, [this, &unonneg, n, m] {
        hineq_.resize(m);
        for (int i = 0; i < m; ++i) hineq_[i] = S_.rL[i] != S_.rU[i];
        F_.nonneg = nonneg_bounds<float>(S_.l, S_.u, n);
        D_.nonneg = nonneg_bounds<double>(S_.l, S_.u, n);
        unonneg = nonneg_bounds<double>(U_.l, U_.u, n);
    });
    try {
        setup_matrices(mixed, true, true);
    } catch (...) {
        init.wait();
        kinds.wait();
        throw;
    }
    init.get();
    kinds.get();
    const double t_mat = seconds_since(t0);
    // one allocation for every buffer: count, allocate, place
    take_buffers(arena_, mixed);
    arena_.commit();
    take_buffers(arena_, mixed);
    // uploads from host arrays that outlive the setup (no staging copies, one
    // synchronisation at the end); the FP32 vectors are rounded on the device
    auto put = [&](const DevBuf& d, const auto* h, size_t len) {
        if (len > 0) SUPR_HIP(hipMemcpyAsync(d.as<void>(), h, sizeof(*h) * len, hipMemcpyHostToDevice, s));
    };
    // vectors whose entries are all one value are filled on the device instead
    // of uploaded (x >= 0 LPs: x0 = clip(0, +0, +inf) = +0, unscaled l = +0,
    // u = +inf): 24 MB less over PCIe on the 1M-variable LP
    auto fill = [&](const DevBuf& d, double v, size_t len) {
        if (len > 0) hipLaunchKernelGGL(k_fill, dim3(grid_elems((long)len)), dim3(kBlock), 0, s, (int64_t)len, v, d.as<double>());
    };
    put(ineq_, hineq_.data(), m);
    if (D_.nonneg) fill(x0d_, 0.0, n);
    else put(x0d_, x0_.data(), n);
    put(D_.c, S_.c, n);
    put(D_.b, S_.rL, m);
    if (!D_.nonneg) { put(D_.l, S_.l, n); put(D_.u, S_.u, n); }
    if (mixed) {
        auto f32 = [&](const DevBuf& src, const DevBuf& dst, size_t len) {
            hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_elems((long)len)), dim3(kBlock), 0, s,
                               (int64_t)len, src.as<const double>(), dst.as<float>());
        };
        f32(D_.c, F_.c, n);
        f32(D_.b, F_.b, m);
        // F_ keeps l, u only when D_ does: +0 / +inf in FP64 stay so in FP32
        if (!F_.nonneg) { f32(D_.l, F_.l, n); f32(D_.u, F_.u, n); }
        SUPR_HIP(hipGetLastError());
    }
    const double t_iter = seconds_since(t0);
    put(kc_, U_.c, n);
    if (unonneg) {
        fill(kl_, 0.0, n);
        fill(ku_, std::numeric_limits<double>::infinity(), n);
    } else {
        put(kl_, U_.l, n); put(ku_, U_.u, n);
    }
    SUPR_HIP(hipGetLastError());
    put(krL_, U_.rL, m); put(krU_, U_.rU, m);
    put(kDr_, Dr_, m); put(kDc_, Dc_, n);
    K_.c = kc_.as<double>(); K_.l = kl_.as<double>(); K_.u = ku_.as<double>();
    K_.rL = krL_.as<double>(); K_.rU = krU_.as<double>();
    K_.Dr = kDr_.as<double>(); K_.Dc = kDc_.as<double>();
    K_.xu = xu_.as<double>(); K_.xu2 = xu2_.as<double>(); K_.xavg = xavg_.as<double>();
    K_.yu = yu_.as<double>(); K_.yu2 = yu2_.as<double>(); K_.yavg = yavg_.as<double>();
    K_.best_x = bestx_.as<double>(); K_.best_y = besty_.as<double>();
    K_.aty_avg = nullptr; K_.ax_avg = axavg_.as<double>();
    double* p = parts_.as<double>();
    part_n_ = p; p += kMaxBlocks;
    part_m_ = p; p += 2 * kMaxBlocks;
    part_kn_ = p; p += 8 * kMaxBlocks;
    part_km_ = p; p += 6 * kMaxBlocks;
    part_rn_ = p; p += kMaxBlocks;
    part_rm_ = p; p += kMaxBlocks;
    dx2_slot_ = p;
    // host-coherent pinned copy of the state, written by the last kernel of
    // every graph replay and polled by the host (no copy on the stream)
    SUPR_HIP(hipHostMalloc((void**)&mirror_, sizeof(DevState), hipHostMallocCoherent | hipHostMallocMapped));
    std::memset(mirror_, 0, sizeof(DevState));
    SUPR_HIP(hipHostGetDevicePointer((void**)&mirror_dev_, mirror_, 0));
    const double t_kkt = seconds_since(t0);
    rebuild_ell_values();

    // SpMV schedules per iterate precision: given, from the shape rule (0),
    // or timed on scratch vectors (-1)
    auto tune = [&](DevMatrix& M, int fixed, auto view, auto* xv, auto* yv) {
        using V = std::remove_pointer_t<decltype(xv)>;
        if (fixed > 0) M.lanes = fixed;
        else if (fixed == 0 || M.ell_w > 0) M.lanes = rule_lanes(M, sizeof(V) == 8);
        else autotune_lanes(M, view, xv, yv, s);
        return M.lanes;
    };
    if (!valid_lanes(o_.spmv_lanes_A <= 0 ? 1 : o_.spmv_lanes_A) ||
        !valid_lanes(o_.spmv_lanes_AT <= 0 ? 1 : o_.spmv_lanes_AT))
        throw BadInput("bad lanes-per-row option");
    if (mixed) {
        lanesA_[0] = tune(A_, o_.spmv_lanes_A, A_.f32(), F_.x[1].as<float>(), F_.ax[1].as<float>());
        lanesAT_[0] = tune(AT_, o_.spmv_lanes_AT, AT_.f32(), F_.y[1].as<float>(), F_.xsum.as<float>());
    }
    lanesA_[1] = tune(A_, o_.spmv_lanes_A, A_.f64(), D_.x[1].as<double>(), D_.ax[1].as<double>());
    lanesAT_[1] = tune(AT_, o_.spmv_lanes_AT, AT_.f64(), D_.y[1].as<double>(), D_.xsum.as<double>());
    SUPR_HIP(hipStreamSynchronize(s));
    if (o_.verbose)
        std::fprintf(stderr, "[suprema] setup: matrices (host passes beside) %.2f ms, iterates %.2f ms, KKT vectors %.2f ms, "
                     "lanes + sync %.2f ms (A %d/%d, AT %d/%d; %.1f MB in one allocation)\n", 1e3 * t_mat,
                     1e3 * (t_iter - t_mat), 1e3 * (t_kkt - t_iter), 1e3 * (seconds_since(t0) - t_kkt),
                     lanesA_[0], lanesA_[1], lanesAT_[0], lanesAT_[1], arena_.bytes() / 1e6);
}

// A^T in ELL with FP64 values that are all (Dr_i a) Dc_j for one unscaled
// value a (transportation, assignment and other matrices whose entries share
// one value): the FP64 primal half forms each value from Dc_j (one per row of
// A^T) and Dr_i a (gathered from an m-vector that stays in cache) instead of
// reading 8 bytes per slot.  Used only if a check on the device finds every
// stored value equal, bit for bit, to the product the kernel forms, so the
// iteration is unchanged.  Not in the plain mode (the reference's design);
// SUPR_ELL_VALUES=stored turns it off.
void Solver::rebuild_ell_values() {
    if (o_.mode == SUPR_MODE_PLAIN_FP64 || AT_.ell_w <= 0 || AT_.nnz <= 0 || m_ <= 0) return;
    if (const char* e = std::getenv("SUPR_ELL_VALUES"); e && std::strcmp(e, "stored") == 0) return;
    hipStream_t s = stream_.s;
    double a = 0.0;   // the unscaled value of A^T's first slot
    SUPR_HIP(hipMemcpyAsync(&a, AT_.vu.as<void>(), sizeof(double), hipMemcpyDeviceToHost, s));
    SUPR_HIP(hipStreamSynchronize(s));
    std::vector<double> cs((size_t)m_);
    for (int i = 0; i < m_; ++i) cs[(size_t)i] = Dr_[i] * a;
    DevBuf c = upload(cs, s);
    DevBuf bad = device_vec<uint32_t>(1);
    SUPR_HIP(hipMemsetAsync(bad.as<void>(), 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_ell_rebuild_check, dim3(grid_elems((long)AT_.nnz)), dim3(kBlock), 0, s, AT_.f64(),
                       c.as<const double>(), kDc_.as<const double>(), bad.as<uint32_t>());
    SUPR_HIP(hipGetLastError());
    uint32_t hb = 1;
    SUPR_HIP(hipMemcpyAsync(&hb, bad.as<void>(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    SUPR_HIP(hipStreamSynchronize(s));
    if (hb != 0) return;
    ellc_ = std::move(c);
    AT_.csc = ellc_.as<double>();
    AT_.rsc = kDc_.as<double>();
}

DevState Solver::initial_state(bool mixed) const {
    DevState st;
    std::memset(&st, 0, sizeof(st));
    st.eta = eta0_;
    st.omega = omega0_;
    st.merit_last_restart = INFINITY;
    st.merit_prev_cand = INFINITY;
    st.best_merit = INFINITY;
    st.stall_ref = INFINITY;
    st.tol = o_.tol;
    double sw = o_.switch_tol;
    if (sw <= 0) sw = o_.tol < 1e-5 ? 1e-5 : 0.0;   // FP32 reaches ~1e-5; looser tolerances finish in FP32
    st.switch_tol = sw;
    st.stall_checks = o_.stall_checks > 0 ? o_.stall_checks : 6;
    st.denom_p = 1.0 + std::max(finite_inf_norm(U_.rL, m_), finite_inf_norm(U_.rU, m_));
    st.denom_d = 1.0 + finite_inf_norm(U_.c, n_);
    st.obj_offset = U_.obj_offset;
    st.max_iter = o_.max_iterations;
    st.check_every = o_.check_every;
    st.active = o_.max_iterations > 0 ? 1 : 0;
    st.need_check = o_.max_iterations > 0 ? 0 : 1;
    st.need_aty = 1;
    st.phase = mixed ? 0 : 1;
    return st;
}

template <typename T> MatView<T> iter_view(const DevMatrix& M);
template <> MatView<float> iter_view<float>(const DevMatrix& M) { return M.f32(); }
template <> MatView<double> iter_view<double>(const DevMatrix& M) { return M.f64_iter(); }

// A^T's view for the primal launches: two adjacent rows per lane
// (MatView::pairs) where A^T is ELL with one lane per row, an even row count
// and more rows than one pass of the capped grid holds.  Each row's values
// are the same bits either way; the step statistic's partial sums group the
// rows differently, so the trajectory differs in the last bits.
template <typename T> MatView<T> primal_view(const DevMatrix& AT, bool pairs_on) {
    MatView<T> v = iter_view<T>(AT);
    v.pairs = pairs_on && AT.ell_w > 0 && AT.lanes == 1 && AT.rows % 2 == 0 &&
              (long)AT.rows > (long)kMaxBlocks * kBlock;
    return v;
}

// The iteration kernels, in their one-pass form (matrix slots and epilogue
// operands loaded before the step decision) when pre is set and the launch
// covers every row in one pass.
// early: FP64 row pairs in one pass of the grid take k_primal_pairs (the
// pair's first index slots loaded before the decision prologue; the same bits
// as k_primal's row-pair path, SUPR_EARLY_SLOTS=off)
template <int L, typename T>
void launch_primal(bool pre, bool early, int grid, hipStream_t s, DevState* S, int p, const MatView<T>& AT,
                   const IterBufs<T>& B, const double* part_m, int nb_m, const double* dx2, double* part_n) {
    if constexpr (sizeof(T) == 8 && L == 1) {
        if (early && AT.pairs && (long)AT.rows <= 2L * grid * kBlock) {
            hipLaunchKernelGGL((k_primal_pairs<T>), dim3(grid), dim3(kBlock), 0, s, S, p, AT, B, part_m, nb_m, dx2, part_n);
            return;
        }
    }
    if (pre && one_pass_rows<L>(AT.rows))
        hipLaunchKernelGGL((k_primal<L, T, true>), dim3(grid), dim3(kBlock), 0, s, S, p, AT, B, part_m, nb_m, dx2, part_n);
    else
        hipLaunchKernelGGL((k_primal<L, T, false>), dim3(grid), dim3(kBlock), 0, s, S, p, AT, B, part_m, nb_m, dx2, part_n);
}
// pre: 0 the grid-stride dual, 1 the one-pass form, 2 the one-pass form
// over long CSR rows (row_dot_flat, row_dot_diff_flat)
template <int L, typename T>
void launch_dual(int pre, int grid, hipStream_t s, DevState* S, int p1, const MatView<T>& A,
                 const IterBufs<T>& B, const double* part_n, int nb_n, double* dx2, double* part_m) {
    if (pre == 2 && one_pass_rows<L>(A.rows) && A.ell_w == 0) {
        hipLaunchKernelGGL((k_dual<L, T, 2>), dim3(grid), dim3(kBlock), 0, s, S, p1, A, B, part_n, nb_n, dx2, part_m);
    } else if (pre != 0 && one_pass_rows<L>(A.rows)) {
        hipLaunchKernelGGL((k_dual<L, T, 1>), dim3(grid), dim3(kBlock), 0, s, S, p1, A, B, part_n, nb_n, dx2, part_m);
    } else {
        hipLaunchKernelGGL((k_dual<L, T, 0>), dim3(grid), dim3(kBlock), 0, s, S, p1, A, B, part_n, nb_n, dx2, part_m);
    }
}

template <typename T>
void Solver::enqueue_attempts(const Iter<T>& I, int count) {
    hipStream_t s = stream_.s;
    DevState* S = state_.as<DevState>();
    const IterBufs<T> B = I.view(ineq_.as<uint8_t>());
    const MatView<T> A = iter_view<T>(A_), AT = primal_view<T>(AT_, row_pairs_);
    const int gn = grid_rows(n_, AT_.lanes), gm = grid_rows(m_, A_.lanes);
    const bool pn = one_pass(AT_);
    const int pm = one_pass_dual(A_);
    for (int k = 0; k < count; ++k) {
        const int p = k & 1;
        SUPR_LANES(AT_.lanes, (launch_primal<LL, T>(pn, early_slots_, gn, s, S, p, AT, B, part_m_, gm, dx2_slot_, part_n_)));
        SUPR_LANES(A_.lanes, (launch_dual<LL, T>(pm, gm, s, S, p ^ 1, A, B, part_n_, gn, dx2_slot_, part_m_)));
    }
    SUPR_HIP(hipGetLastError());
}

// A x of the current iterate with the plain SpMV (start-up and the FP32 ->
// FP64 switch), on the stream; outside any graph.  (A^T y is formed by every
// primal launch.)
template <typename T>
void Solver::set_products(const Iter<T>& I, const MatView<T>& A) {
    launch_spmv(A_, A, I.x[0].template as<T>(), I.ax[0].template as<T>(), stream_.s);
}

template <int LA>
void launch_check_kkt(int lat, hipStream_t s, const DevState* S, const MatView<double>& AuT,
                      const MatView<double>& Au, int gkn, int gkm, const KktBufs& K, double* pkn,
                      double* pkm) {
    SUPR_LANES(lat, hipLaunchKernelGGL((k_check_kkt<LA, LL>), dim3(gkn + gkm), dim3(kBlock), 0, s, S, AuT,
                                       Au, gkn, K, pkn, pkm));
}

template <typename T>
void Solver::enqueue_check(const Iter<T>& I) {
    hipStream_t s = stream_.s;
    DevState* S = state_.as<DevState>();
    const IterBufs<T> B = I.view(ineq_.as<uint8_t>());
    const int en = grid_check_elems(n_), em = grid_check_elems(m_);
    const int gm = grid_rows(m_, A_.lanes);
    hipLaunchKernelGGL((k_check_prep<T>), dim3(en + em), dim3(kBlock), 0, s, S, n_, m_, en, B, K_, part_m_,
                       gm, dx2_slot_);
    const int gkn = grid_check_rows(n_, AT_.lanes), gkm = grid_check_rows(m_, A_.lanes);
    const MatView<double> AuT = AT_.unscaled(), Au = A_.unscaled();
    SUPR_LANES(A_.lanes, launch_check_kkt<LL>(AT_.lanes, s, S, AuT, Au, gkn, gkm, K_, part_kn_, part_km_));
    hipLaunchKernelGGL(k_kkt_fin, dim3(1), dim3(kBlock), 0, s, S, part_kn_, gkn, part_km_, gkm);
    hipLaunchKernelGGL((k_check_restart<T>), dim3(en + em), dim3(kBlock), 0, s, S, n_, m_, en, B, K_,
                       part_rn_, part_rm_);
    hipLaunchKernelGGL(k_restart_fin, dim3(1), dim3(kBlock), 0, s, S, part_rn_, en, part_rm_, em, mirror_dev_);
    SUPR_HIP(hipGetLastError());
}

template <typename T>
hipGraphExec_t Solver::capture_graph(const Iter<T>& I, int attempts, int units) {
    hipStream_t s = stream_.s;
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    SUPR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        for (int u = 0; u < units; ++u) {
            enqueue_attempts(I, attempts);
            enqueue_check(I);
        }
    } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    SUPR_HIP(hipStreamEndCapture(s, &g));
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    SUPR_HIP(e);
    return exec;
}

// Launch plan of the direct-launch loop (auto mode, large LPs).  A KKT
// check becomes due when `accepted` reaches the next multiple of check_every
// (or `attempts` reaches max_iterations); until then a check group is a
// no-op, and once it is due the attempts are no-ops until a check group
// runs.  Instead of a check group after every 16 attempts, the host places
// one where the check is expected to become due (the attempts the accepted
// steps need at the rejection rate seen so far), then one more after a
// short unit in case more attempts were rejected, and launches the attempts
// between them in units without checks.  It plans from the state the last
// check group left in the host-coherent copy (which group it was is the
// state's `groups` count), so it never waits for the device.  The device
// follows the same state machine whatever the plan: only the number of
// no-op launches changes, never the bits.
struct CheckPlan {
    int check_every = 64, max_iter = 0;
    int64_t launched = 0;                 // attempts launched in this loop
    std::vector<int64_t> group_at;        // `launched` after each check group of this loop
    int32_t base_groups = 0;              // state.groups when the loop started
    bool slack_due = false;               // the last unit ended at an expected check
    static constexpr int kMaxUnit = 32;   // attempts per unit without a check group
    static constexpr int kSlack = 4;      // attempts in the unit after an expected check
    static constexpr int kRateSpan = 128; // attempts the rejection rate is taken over
    // (attempts, accepted) of each state the host has seen, one per check
    // group count
    std::vector<std::pair<int32_t, int32_t>> seen;
    int32_t seen_groups = -1;

    // The rejection rate of the last kRateSpan or more attempts (the loop's
    // whole rate until it has that many).  The rate drifts down as the step
    // size settles: on the 1M-variable LP the whole loop's 5% placed the late
    // checks' groups 4 attempts after the check came due (4 no-op attempts
    // per check), where the recent rate is 0.
    double rejection(const DevState& hs) {
        if (hs.groups != seen_groups) {
            seen.push_back({hs.attempts, hs.accepted});
            seen_groups = hs.groups;
        }
        int32_t a0 = 0, c0 = 0;
        for (size_t j = seen.size(); j-- > 0;)
            if (seen[j].first <= hs.attempts - kRateSpan) {
                a0 = seen[j].first;
                c0 = seen[j].second;
                break;
            }
        const int32_t da = hs.attempts - a0;
        return da > 0 ? std::min(0.5, (double)(da - (hs.accepted - c0)) / da) : 0.0;
    }

    // The next unit from the last known state: attempts (even) and whether
    // a check group follows them.
    std::pair<int, bool> next(const DevState& hs) {
        if (slack_due) {
            slack_due = false;
            return {kSlack, true};
        }
        const int64_t jp = std::max<int64_t>(0, hs.groups - base_groups);
        const int64_t at = jp == 0 ? 0 : group_at[(size_t)std::min<int64_t>(jp, (int64_t)group_at.size()) - 1];
        const double rej = rejection(hs);
        const int64_t ce = check_every;
        const int64_t due_acc = hs.need_check ? hs.accepted : ((int64_t)hs.accepted / ce + 1) * ce;
        int64_t target = at + (int64_t)std::ceil((double)(due_acc - hs.accepted) / (1.0 - rej));
        target = std::min<int64_t>(target, at + std::max<int64_t>(0, (int64_t)max_iter - hs.attempts));
        // check groups already launched at or after the expected point are
        // taken to have run that check; expect the next one a check interval on
        const int64_t interval = (int64_t)std::ceil((double)ce / (1.0 - rej));
        for (size_t j = (size_t)jp; j < group_at.size(); ++j)
            if (group_at[j] >= target) target = group_at[j] + interval;
        int64_t k = target - launched;
        if (k <= 0) return {2, true};
        if (k > kMaxUnit) return {kMaxUnit, false};
        k += k & 1;
        slack_due = true;
        return {(int)k, true};
    }
    void launched_unit(int k, bool check) {
        launched += k;
        if (check) group_at.push_back(launched);
    }
};

// Replays the graph until the device reports done / a precision switch, or
// the stop callback fires.  A few replays stay in flight; after the oldest
// one's event the host reads the host-coherent state copy the replays'
// last kernel writes (it may already be a newer replay's: the flags only
// move forward, and the final state is read back with a synchronous copy).
template <typename T>
int Solver::run_graph_loop(const Iter<T>& I, int prec, supr_stop_fn stop, void* ctx,
                           DevState& hs, bool& stopped) {
    hipStream_t s = stream_.s;
    // A poll unit is K times (G attempts + the check group).
    //   graph_attempts > 0: one graph replay of G = graph_attempts, K = 1;
    //   graph_attempts < 0: the same kernels launched on the stream directly,
    //                       G = -graph_attempts, K = 1;
    //   0 (auto): matrices of 2^20+ nonzeros launch directly (G = 16): each
    //     launch runs several us longer than the host takes to enqueue it,
    //     and direct launches skip the graph-launch boundary (on MI355X about
    //     1 ms less per solve of the 1M-variable LP than 16-attempt replays);
    //     smaller ones replay graphs of 8 units of 16 attempts (the random
    //     2,000 x 5,000 LP: 10.3 against 10.8 us per attempt at one unit, and
    //     12.1 launched directly, where the host cannot keep ahead).
    //   0 (auto), 2^20+ nonzeros: launched directly, planned by CheckPlan
    //     (units of up to 32 attempts, check groups where checks are expected).
    const bool autodirect = o_.graph_attempts == 0 && A_.nnz >= (1L << 20);
    const bool direct = o_.graph_attempts < 0 || autodirect;
    int G = o_.graph_attempts != 0 ? std::abs(o_.graph_attempts) : 16;
    G += G & 1;   // the state double buffer needs an even number of attempts per unit
    const int K = o_.graph_attempts == 0 && !direct ? 8 : 1;
    if (!direct && !graph_[prec]) graph_[prec] = capture_graph(I, G, K);
    hipGraphExec_t exec = graph_[prec];
    CheckPlan plan;
    plan.check_every = o_.check_every;
    plan.max_iter = o_.max_iterations;
    plan.base_groups = hs.groups;   // hs: the state the loop starts from
    // The host-coherent copy holds the last check group's state, which may be
    // the previous loop's (done, or a precision switch); polls before this
    // loop's first check group must see this loop's start instead.  Nothing
    // on the device writes it now: the previous loop ended with a stream sync.
    std::memcpy((void*)mirror_, &hs, sizeof(DevState));
    std::atomic_thread_fence(std::memory_order_release);
    // units in flight: a replay takes far longer than the host's poll; the
    // planned units can be short, so one more is kept queued
    const int R = autodirect ? 3 : 2;
    hipEvent_t ev[3] = {};
    int replays = 0;
    try {
        for (int r = 0; r < R; ++r) SUPR_HIP(hipEventCreateWithFlags(&ev[r], hipEventDisableTiming));
        int last_attempts = -1, last_checks = -1, last_groups = hs.groups, idle = 0;
        auto last_log = Clock::now();
        while (true) {
            if (replays >= R) {
                const int slot = replays % R;
                SUPR_HIP(hipEventSynchronize(ev[slot]));
                std::atomic_thread_fence(std::memory_order_acquire);
                std::memcpy(&hs, (const void*)mirror_, sizeof(DevState));
                if (hs.done || hs.switch_req) break;
                if (stop && stop(ctx)) { stopped = true; break; }
                // A check group always advances attempts or checks; a loop
                // that stops moving is a logic error -- fail instead of spinning.
                if (hs.groups != last_groups) {
                    if (hs.attempts == last_attempts && hs.checks == last_checks) {
                        if (++idle > 64) throw std::runtime_error("device loop made no progress");
                    } else {
                        idle = 0;
                    }
                    last_groups = hs.groups;
                }
                last_attempts = hs.attempts;
                last_checks = hs.checks;
                if (o_.verbose >= 2 && seconds_since(last_log) >= 5.0) {
                    std::fprintf(stderr, "[suprema] %s loop: attempts=%d accepted=%d checks=%d restarts=%d merit=%.3e best=%.3e\n",
                                 prec ? "fp64" : "fp32", hs.attempts, hs.accepted, hs.checks, hs.restarts,
                                 hs.cand_merit, hs.best_merit);
                    last_log = Clock::now();
                }
            }
            const int slot = replays % R;
            if (autodirect) {
                const auto [k, check] = plan.next(hs);
                enqueue_attempts(I, k);
                if (check) enqueue_check(I);
                plan.launched_unit(k, check);
            } else if (direct) {
                for (int u = 0; u < K; ++u) {
                    enqueue_attempts(I, G);
                    enqueue_check(I);
                }
            } else {
                SUPR_HIP(hipGraphLaunch(exec, s));
            }
            SUPR_HIP(hipEventRecord(ev[slot], s));
            ++replays;
        }
        SUPR_HIP(hipStreamSynchronize(s));
        SUPR_HIP(hipMemcpy(&hs, state_.as<void>(), sizeof(DevState), hipMemcpyDeviceToHost));
    } catch (...) {
        (void)hipStreamSynchronize(s);
        for (auto& e : ev) if (e) (void)hipEventDestroy(e);
        throw;
    }
    for (auto& e : ev) if (e) (void)hipEventDestroy(e);
    return replays;
}

void Solver::run_device_resident(bool mixed, supr_stop_fn stop, void* ctx, double* x_out,
                                 double* y_out, supr_pdlp_info* info) {
    hipStream_t s = stream_.s;
    const auto t0 = Clock::now();
    const int prec0 = mixed ? 0 : 1;
    use_lanes(prec0);
    // the cold start, the state and A x0 are enqueued without a host wait:
    // the loop's first launches follow them on the stream
    const DevState st0 = initial_state(mixed);
    if (mixed) reset_iter(F_, &st0);
    else reset_iter(D_, &st0);
    if (mixed) set_products(F_, A_.f32());
    else set_products(D_, A_.f64());
    const auto t1 = Clock::now();

    DevState hs = st0;   // the state the loop starts from (its launch plan reads it)
    bool stopped = false;
    int fp32_attempts = 0;
    if (mixed) {
        run_graph_loop(F_, 0, stop, ctx, hs, stopped);
        fp32_attempts = hs.attempts;
        if (o_.verbose)
            std::fprintf(stderr, "[suprema] fp32 loop: attempts=%d restarts=%d checks=%d merit=%.3e best=%.3e switch=%d done=%d\n",
                         hs.attempts, hs.restarts, hs.checks, hs.cand_merit, hs.best_merit, hs.switch_req, hs.done);
        if (hs.switch_req && !stopped) {
            // Continue in FP64 from the FP32 loop's candidate, as a restart.
            const IterBufs<float> Fv = F_.view(ineq_.as<uint8_t>());
            const IterBufs<double> Dv = D_.view(ineq_.as<uint8_t>());
            hipLaunchKernelGGL(k_switch_n, dim3(grid_elems(n_)), dim3(kBlock), 0, s, state_.as<DevState>(), n_, Fv, Dv, K_);
            hipLaunchKernelGGL(k_switch_m, dim3(grid_elems(m_)), dim3(kBlock), 0, s, state_.as<DevState>(), m_, Fv, Dv, K_);
            SUPR_HIP(hipGetLastError());
            switch_to_fp64(hs);
            const DevState both[2] = {hs, hs};
            SUPR_HIP(hipMemcpyAsync(state_.as<void>(), both, sizeof(both), hipMemcpyHostToDevice, s));
            use_lanes(1);
            set_products(D_, A_.f64());
            SUPR_HIP(hipStreamSynchronize(s));
            run_graph_loop(D_, 1, stop, ctx, hs, stopped);
        }
    } else {
        run_graph_loop(D_, 1, stop, ctx, hs, stopped);
    }
    const double t_solve = seconds_since(t1);
    const double t_prep = std::chrono::duration<double>(t1 - t0).count();

    bool has_point = hs.has_best != 0;
    // c'x + offset of the returned point: the KKT check formed it on the
    // device (FP64, fixed order) when it kept that point
    double obj = has_point ? hs.best_obj : U_.obj_offset;
    if (has_point) {
        SUPR_HIP(hipMemcpy(x_out, bestx_.as<void>(), sizeof(double) * n_, hipMemcpyDeviceToHost));
        SUPR_HIP(hipMemcpy(y_out, besty_.as<void>(), sizeof(double) * m_, hipMemcpyDeviceToHost));
    } else {
        // No check ran yet (stopped early): assess the current point on the host.
        std::vector<double> xh(n_), yh(m_);
        if (hs.phase == 0) {
            std::vector<float> xf(n_), yf(m_);
            SUPR_HIP(hipMemcpy(xf.data(), F_.x[hs.cur].as<void>(), sizeof(float) * n_, hipMemcpyDeviceToHost));
            SUPR_HIP(hipMemcpy(yf.data(), F_.y[hs.cur].as<void>(), sizeof(float) * m_, hipMemcpyDeviceToHost));
            for (int j = 0; j < n_; ++j) xh[j] = Dc_[j] * (double)xf[j];
            for (int i = 0; i < m_; ++i) yh[i] = Dr_[i] * (double)yf[i];
        } else {
            SUPR_HIP(hipMemcpy(xh.data(), D_.x[hs.cur].as<void>(), sizeof(double) * n_, hipMemcpyDeviceToHost));
            SUPR_HIP(hipMemcpy(yh.data(), D_.y[hs.cur].as<void>(), sizeof(double) * m_, hipMemcpyDeviceToHost));
            for (int j = 0; j < n_; ++j) xh[j] *= Dc_[j];
            for (int i = 0; i < m_; ++i) yh[i] *= Dr_[i];
        }
        double rep[7];
        const supr_lp& Uh = unscaled_with_matrix();
        const Csr Au = build_csr(Uh, Uh.val, false), AuT = build_csr(Uh, Uh.val, true);
        verify_csr(Uh, Au, AuT, xh.data(), yh.data(), rep);
        if (std::isfinite(merit_of(rep))) {
            has_point = true;
            std::copy(xh.begin(), xh.end(), x_out);
            std::copy(yh.begin(), yh.end(), y_out);
            hs.best_eps[0] = rep[0]; hs.best_eps[1] = rep[1]; hs.best_eps[2] = rep[2];
            obj = rep[3];
        }
    }
    info->iterations = hs.attempts;
    info->restarts = hs.restarts;
    info->converged = hs.done == 1 ? 1 : 0;
    info->has_point = has_point ? 1 : 0;
    info->eps_P = hs.best_eps[0]; info->eps_D = hs.best_eps[1]; info->eps_G = hs.best_eps[2];
    info->objective = obj;
    const double t_out = seconds_since(t1) - t_solve;
    info->iterations_fp32 = fp32_attempts;
    info->checks = hs.checks;
    info->lanes_A = A_.lanes; info->lanes_AT = AT_.lanes;
    info->seconds_solve = t_solve;
    if (o_.verbose)
        std::fprintf(stderr, "[suprema] done: attempts=%d (fp32 %d) accepted=%d restarts=%d checks=%d eps=%.2e/%.2e/%.2e conv=%d solve=%.4fs (start %.2f ms, loop, result %.2f ms) setup=%.4fs lanes A=%d AT=%d ell A=%d AT=%d\n",
                     hs.attempts, fp32_attempts, hs.accepted, hs.restarts, hs.checks, info->eps_P, info->eps_D, info->eps_G,
                     info->converged, t_solve, 1e3 * t_prep, 1e3 * t_out, setup_s_, A_.lanes, AT_.lanes, A_.ell_w, AT_.ell_w);
}

// Standalone timing of the two iteration kernels at their real shapes (the
// kernels the graph replays), for the roofline figure:
//   out[0] ms per k_primal, out[1] ms per k_dual (FP32), out[2], out[3] the
//   same in FP64, out[4] lanes A, out[5] lanes A^T, out[6] ELL width of A,
//   out[7] ELL width of A^T, out[8] 1 if the FP64 primal half forms A^T's
//   values from factors (rebuild_ell_values).
// The state is set so every launch does a full accepted step's work (A^T y,
// average terms, projections, A x) and no decision is pending: k_primal
// reads S[0] and writes S[1], k_dual reads S[1], so repeated launches see the
// same state.
void Solver::bench_kernels(int reps, double* out) {
    hipStream_t s = stream_.s;
    if (o_.mode != SUPR_MODE_MIXED || !ready_) throw BadInput("bench_kernels needs a set-up mixed-mode engine");
    use_lanes(0);
    reset_iter(F_);
    set_products(F_, A_.f32());
    DevState st = initial_state(true);
    st.check_every = std::numeric_limits<int32_t>::max();
    st.max_iter = std::numeric_limits<int32_t>::max();
    st.pending = 1;
    st.eta_acc = st.eta;
    st.need_aty = 1;
    st.decide_pending = 0;
    hipEvent_t e0, e1;
    SUPR_HIP(hipEventCreate(&e0));
    SUPR_HIP(hipEventCreate(&e1));
    auto time_kernel = [&](auto&& launch) {
        for (int r = 0; r < 3; ++r) launch();
        SUPR_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) launch();
        SUPR_HIP(hipEventRecord(e1, s));
        SUPR_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SUPR_HIP(hipEventElapsedTime(&ms, e0, e1));
        return (double)ms / reps;
    };
    auto run = [&](auto& I, auto tag, int slot) {
        using T = decltype(tag);
        const IterBufs<T> B = I.view(ineq_.as<uint8_t>());
        const MatView<T> A = iter_view<T>(A_), AT = primal_view<T>(AT_, row_pairs_);
        const int gn = grid_rows(n_, AT_.lanes), gm = grid_rows(m_, A_.lanes);
        DevState* S = state_.as<DevState>();
        const DevState both[2] = {st, st};
        SUPR_HIP(hipMemcpyAsync(S, both, sizeof(both), hipMemcpyHostToDevice, s));
        out[slot] = time_kernel([&] {
            SUPR_LANES(AT_.lanes, (launch_primal<LL, T>(one_pass(AT_), early_slots_, gn, s, S, 0, AT, B, part_m_, gm, dx2_slot_, part_n_)));
        });
        out[slot + 1] = time_kernel([&] {
            SUPR_LANES(A_.lanes, (launch_dual<LL, T>(one_pass_dual(A_), gm, s, S, 1, A, B, part_n_, gn, dx2_slot_, part_m_)));
        });
        SUPR_HIP(hipGetLastError());
    };
    run(F_, float{}, 0);
    use_lanes(1);
    reset_iter(D_);
    set_products(D_, A_.f64());
    run(D_, double{}, 2);
    SUPR_HIP(hipStreamSynchronize(s));
    out[4] = A_.lanes; out[5] = AT_.lanes; out[6] = A_.ell_w; out[7] = AT_.ell_w;
    out[8] = AT_.csc ? 1.0 : 0.0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

int check_options(const supr_lp* S, const supr_pdlp_options* opt);

int check_solve_args(const supr_lp* U, const supr_lp* S, const double* Dr, const double* Dc,
                     const supr_pdlp_options* opt) {
    // U's indices equal S's (checked below), so only S's are range-checked
    if (!check_lp(U, "unscaled_form", false) || !check_lp(S, "scaled_form")) return SUPR_EINVAL;
    if (U->m != S->m || U->n != S->n || U->nnz != S->nnz) {
        set_error("unscaled and scaled forms differ in shape");
        return SUPR_EINVAL;
    }
    const size_t pat = sizeof(int32_t) * (size_t)S->nnz;
    if (S->nnz > 0 && (std::memcmp(U->row_idx, S->row_idx, pat) != 0 ||
                       std::memcmp(U->col_idx, S->col_idx, pat) != 0)) {
        for (int k = 0; k < S->nnz; ++k)
            if (U->row_idx[k] != S->row_idx[k] || U->col_idx[k] != S->col_idx[k]) {
                set_error("unscaled and scaled forms must share one COO pattern (entry " + std::to_string(k) + ")");
                return SUPR_EINVAL;
            }
    }
    if (!Dr || !Dc) { set_error("null Dr or Dc"); return SUPR_EINVAL; }
    return check_options(S, opt);
}

// Shape and option checks shared by every single-LP entry point.
int check_options(const supr_lp* S, const supr_pdlp_options* opt) {
    // a form without rows (box constraints only) or without nonzeros is solved
    // like any other, as the reference's pdlp_run does (include/pdlp_algo.hpp:
    // 67-75: eta0 = 1 without entries, omega0 = 1 without finite rL)
    if (S->n < 1) {
        set_error("the PDLP form needs at least one column");
        return SUPR_EINVAL;
    }
    if (!opt) { set_error("null options"); return SUPR_EINVAL; }
    if (opt->max_iterations < 0 || opt->check_every < 1 || !(opt->tol >= 0)) {
        set_error("bad options: max_iterations >= 0, check_every >= 1, tol >= 0 required");
        return SUPR_EINVAL;
    }
    if ((opt->spmv_lanes_A > 0 && !valid_lanes(opt->spmv_lanes_A)) ||
        (opt->spmv_lanes_AT > 0 && !valid_lanes(opt->spmv_lanes_AT)) || opt->spmv_lanes_A < -1 ||
        opt->spmv_lanes_AT < -1) {
        set_error("spmv lanes must be 0 (shape rule), -1 (timed autotune), a power of two in [1, 64], or 256");
        return SUPR_EINVAL;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
        set_error("no HIP device visible");
        return SUPR_ENODEV;
    }
    if (opt->device < 0 || opt->device >= count) {
        set_error("device ordinal out of range");
        return SUPR_EINVAL;
    }
    return SUPR_OK;
}

// Arguments of the unscaled-form entry points.
int check_form_args(const supr_lp* form, int32_t ruiz_passes, const supr_pdlp_options* opt) {
    if (!check_lp(form, "form")) return SUPR_EINVAL;
    if (ruiz_passes < 0) { set_error("negative ruiz_passes"); return SUPR_EINVAL; }
    return check_options(form, opt);
}

}  // namespace
}  // namespace supr

extern "C" {

int supr_device_count(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return 0;
    return count;
}

void supr_pdlp_default_options(supr_pdlp_options* opt) {
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->mode = SUPR_MODE_FAST_FP64;   // the reference's arithmetic; MIXED is opt-in
    opt->max_iterations = 200000;
    opt->check_every = 64;
    opt->tol = 1e-6;
}

int supr_solve_pdhg_gpu(const supr_lp* unscaled_form, const supr_lp* scaled_form,
                        const double* Dr, const double* Dc, const supr_pdlp_options* opt,
                        supr_stop_fn stop, void* stop_ctx, double* x_out, double* y_out,
                        supr_pdlp_info* info) {
    using namespace supr;
    const auto t0 = Clock::now();
    int rc = check_solve_args(unscaled_form, scaled_form, Dr, Dc, opt);
    if (rc != SUPR_OK) return rc;
    if (!x_out || !y_out || !info) { set_error("null output"); return SUPR_EINVAL; }
    const double t_args = seconds_since(t0);
    try {
        SUPR_HIP(hipSetDevice(opt->device));
        ColumnRelabel R;
        const bool rl = relabel_columns(*unscaled_form, *scaled_form, Dc, opt->mode, R);
        const double t_relabel = seconds_since(t0);
        double t_run = 0.0;
        {
            Solver solver(rl ? R.U : *unscaled_form, rl ? R.S : *scaled_form, Dr, rl ? R.dc.data() : Dc, *opt,
                          OwnedLp::Borrow{});
            solver.setup();
            std::vector<double> xp(rl ? (size_t)R.S.n : 0, 0.0);
            solver.run(stop, stop_ctx, rl ? xp.data() : x_out, y_out, info);
            if (rl) map_back(R.newpos, xp.data(), x_out, info);
            t_run = seconds_since(t0);
        }
        if (opt->verbose)
            std::fprintf(stderr, "[suprema] one-shot call: argument checks %.2f ms, relabel rule %.2f ms, "
                         "setup + solve %.2f ms, teardown %.2f ms\n", 1e3 * t_args,
                         1e3 * (t_relabel - t_args), 1e3 * (t_run - t_relabel), 1e3 * (seconds_since(t0) - t_run));
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}

struct supr_engine {
    int device;
    std::unique_ptr<supr::Solver> solver;
    // columns relabelled for the solver (empty: as given): x is mapped back
    // through newpos from xp
    std::vector<int32_t> newpos;
    std::vector<double> xp;
};

int supr_solve_form_gpu(const supr_lp* form, const char* integer, int32_t ruiz_passes,
                        const supr_pdlp_options* opt, supr_stop_fn stop, void* stop_ctx,
                        double* x_out, double* y_out, supr_pdlp_info* info) {
    using namespace supr;
    int rc = check_form_args(form, ruiz_passes, opt);
    if (rc != SUPR_OK) return rc;
    if (!x_out || !y_out || !info) { set_error("null output"); return SUPR_EINVAL; }
    try {
        SUPR_HIP(hipSetDevice(opt->device));
        Solver solver(*form, integer, ruiz_passes, *opt);
        solver.setup();
        solver.run(stop, stop_ctx, x_out, y_out, info);
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}

// The reference's solve_mps() PDLP stage after presolve (src/solve.cpp:194-
// 224), with every O(nnz) step on the device: build_pdlp_form, scale(), the
// solve, then extract_solution, z = c - A^T y and verify() on the original
// LP; while the verifier misses tol, solve again with a 10x tighter internal
// tolerance (three attempts at most).
int supr_solve_lp_gpu(const supr_lp* lp, const supr_pdlp_options* opt, supr_stop_fn stop,
                      void* stop_ctx, double* x_out, double* y_out, double* z_out,
                      supr_pdlp_info* info, double* report7) {
    using namespace supr;
    if (!check_lp(lp, "supr_solve_lp_gpu")) return SUPR_EINVAL;
    if (!opt || !x_out || !y_out || !info || !report7) { set_error("null argument"); return SUPR_EINVAL; }
    if (opt->mode == SUPR_MODE_PLAIN_FP64) {
        set_error("supr_solve_lp_gpu runs the device-resident modes");
        return SUPR_EINVAL;
    }
    try {
        SUPR_HIP(hipSetDevice(opt->device));
        StreamGuard sg;
        hipStream_t s = sg.s;
        const auto t0 = Clock::now();
        DevLp orig(*lp, s);   // the original LP, laid out for the verifier
        DevForm F;
        build_form_on_device(orig, F, s);
        // the form's vectors on the host (sizes, objective, KKT denominators)
        std::vector<double> c2(F.n2), l2(F.n2), u2(F.n2), rL2(F.m2), rU2(F.m2), signs(F.m2);
        std::vector<int32_t> perm(F.m2);
        auto down = [&](void* dst, const DevBuf& src, size_t bytes) {
            if (bytes > 0) SUPR_HIP(hipMemcpyAsync(dst, src.as<void>(), bytes, hipMemcpyDeviceToHost, s));
        };
        down(c2.data(), F.c, 8 * c2.size()); down(l2.data(), F.l, 8 * l2.size());
        down(u2.data(), F.u, 8 * u2.size()); down(rL2.data(), F.rL, 8 * rL2.size());
        down(rU2.data(), F.rU, 8 * rU2.size()); down(signs.data(), F.row_signs, 8 * signs.size());
        down(perm.data(), F.row_perm, 4 * perm.size());
        SUPR_HIP(hipStreamSynchronize(s));
        supr_lp fv{};
        fv.m = F.m2; fv.n = F.n2; fv.nnz = F.nnz2;
        fv.c = c2.data(); fv.l = l2.data(); fv.u = u2.data(); fv.rL = rL2.data(); fv.rU = rU2.data();
        fv.obj_offset = lp->obj_offset;
        if (fv.n < 1) throw BadInput("the PDLP form needs at least one column");
        auto fdev = std::make_unique<DevLp>(F.m2, F.n2, F.nnz2, std::move(F.row), std::move(F.col),
                                            std::move(F.val), std::move(F.c), std::move(F.l),
                                            std::move(F.u), std::move(F.rL), std::move(F.rU), s);
        int rc = check_options(&fv, opt);
        if (rc != SUPR_OK) return rc;
        Solver solver(std::move(fdev), fv, *opt);
        solver.setup();
        const double setup_s = seconds_since(t0);
        std::vector<double> x2(F.n2), y2(F.m2);
        DevBuf dx = device_vec<double>(lp->n), dy = device_vec<double>(lp->m);
        DevBuf dax = device_vec<double>(lp->m), daty = device_vec<double>(lp->n);
        std::vector<double> ax(lp->m), aty(lp->n);
        int total = 0, restarts = 0;
        double inner = opt->tol, solve_s = 0.0;
        supr_pdlp_info step{};
        for (int attempt = 0; attempt < 3; ++attempt, inner *= 0.1) {
            solver.set_tol(inner);
            solver.run(stop, stop_ctx, x2.data(), y2.data(), &step);
            total += step.iterations;
            restarts = step.restarts;
            solve_s += step.seconds_solve;
            // extract_solution (src/pdlp.cpp:84-90)
            std::copy(x2.begin(), x2.begin() + lp->n, x_out);
            std::fill(y_out, y_out + lp->m, 0.0);
            for (int32_t p = 0; p < F.m2; ++p) y_out[perm[p]] = signs[p] * y2[p];
            if (lp->n > 0) SUPR_HIP(hipMemcpyAsync(dx.as<void>(), x_out, 8 * (size_t)lp->n, hipMemcpyHostToDevice, s));
            if (lp->m > 0) SUPR_HIP(hipMemcpyAsync(dy.as<void>(), y_out, 8 * (size_t)lp->m, hipMemcpyHostToDevice, s));
            verify_products_on_device(orig, dx.as<double>(), dy.as<double>(), dax.as<double>(),
                                      daty.as<double>(), s);
            down(ax.data(), dax, 8 * ax.size());
            down(aty.data(), daty, 8 * aty.size());
            SUPR_HIP(hipStreamSynchronize(s));
            verify_products(*lp, x_out, y_out, ax.data(), aty.data(), report7);
            if (z_out)
                for (int32_t j = 0; j < lp->n; ++j) z_out[j] = lp->c[j] - aty[j];
            const double worst = std::max({report7[0], report7[1], report7[2]});
            if (!step.converged || worst <= opt->tol) break;
        }
        *info = step;
        info->iterations = total;
        info->restarts = restarts;
        info->eps_P = report7[0]; info->eps_D = report7[1]; info->eps_G = report7[2];
        info->objective = report7[3];
        info->seconds_setup = setup_s;
        info->seconds_solve = solve_s;
        info->seconds_total = seconds_since(t0);
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}

int supr_engine_create_form(const supr_lp* form, const char* integer, int32_t ruiz_passes,
                            const supr_pdlp_options* opt, supr_engine** out) {
    using namespace supr;
    int rc = check_form_args(form, ruiz_passes, opt);
    if (rc != SUPR_OK) return rc;
    if (!out) { set_error("null engine output"); return SUPR_EINVAL; }
    *out = nullptr;
    try {
        SUPR_HIP(hipSetDevice(opt->device));
        auto e = std::make_unique<supr_engine>();
        e->device = opt->device;
        e->solver = std::make_unique<Solver>(*form, integer, ruiz_passes, *opt);
        e->solver->setup();
        *out = e.release();
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}

int supr_engine_create(const supr_lp* unscaled_form, const supr_lp* scaled_form, const double* Dr,
                       const double* Dc, const supr_pdlp_options* opt, supr_engine** out) {
    using namespace supr;
    int rc = check_solve_args(unscaled_form, scaled_form, Dr, Dc, opt);
    if (rc != SUPR_OK) return rc;
    if (!out) { set_error("null engine output"); return SUPR_EINVAL; }
    *out = nullptr;
    try {
        SUPR_HIP(hipSetDevice(opt->device));
        auto e = std::make_unique<supr_engine>();
        e->device = opt->device;
        ColumnRelabel R;
        if (relabel_columns(*unscaled_form, *scaled_form, Dc, opt->mode, R)) {
            e->solver = std::make_unique<Solver>(R.U, R.S, Dr, R.dc.data(), *opt);
            e->newpos = std::move(R.newpos);
            e->xp.assign(unscaled_form->n, 0.0);
        } else {
            e->solver = std::make_unique<Solver>(*unscaled_form, *scaled_form, Dr, Dc, *opt);
        }
        e->solver->setup();
        *out = e.release();
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}

int supr_engine_solve(supr_engine* engine, supr_stop_fn stop, void* stop_ctx, double* x_out,
                      double* y_out, supr_pdlp_info* info) {
    using namespace supr;
    if (!engine || !engine->solver || !x_out || !y_out || !info) {
        set_error("supr_engine_solve: null argument");
        return SUPR_EINVAL;
    }
    try {
        SUPR_HIP(hipSetDevice(engine->device));
        if (!engine->newpos.empty()) {
            engine->solver->run(stop, stop_ctx, engine->xp.data(), y_out, info);
            map_back(engine->newpos, engine->xp.data(), x_out, info);
            return SUPR_OK;
        }
        engine->solver->run(stop, stop_ctx, x_out, y_out, info);
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}

void supr_engine_destroy(supr_engine* engine) {
    if (!engine) return;
    (void)hipSetDevice(engine->device);
    delete engine;
}

// y = A x (transposed = 0) or A^T x (transposed = 1) with one schedule.
int supr_spmv(const supr_lp* lp, int32_t transposed, int32_t lanes, int32_t fp32,
              const double* x, double* y) {
    using namespace supr;
    if (!check_lp(lp, "supr_spmv")) return SUPR_EINVAL;
    if (!x || !y) { set_error("supr_spmv: null x or y"); return SUPR_EINVAL; }
    if (lanes != 0 && lanes != -1 && !valid_lanes(lanes)) { set_error("supr_spmv: bad lanes"); return SUPR_EINVAL; }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) { set_error("no HIP device visible"); return SUPR_ENODEV; }
    try {
        StreamGuard stream;
        hipStream_t s = stream.s;
        const Csr H = build_csr(*lp, lp->val, transposed != 0);
        DevMatrix M;
        build_dev_matrix(M, H, lp->val, fp32 != 0, fp32 == 0, false, s);
        const int rows = M.rows, cols = M.cols;
        std::vector<double> out(rows);
        if (fp32) {
            std::vector<float> xf(x, x + cols);
            DevBuf dx = upload(xf, s), dy = device_vec<float>(rows);
            if (lanes == 0) M.lanes = rule_lanes(M, false);
            else if (lanes < 0) autotune_lanes(M, M.f32(), dx.as<float>(), dy.as<float>(), s);
            else M.lanes = lanes;
            launch_spmv(M, M.f32(), dx.as<float>(), dy.as<float>(), s);
            std::vector<float> yf(rows);
            SUPR_HIP(hipMemcpyAsync(yf.data(), dy.as<void>(), sizeof(float) * rows, hipMemcpyDeviceToHost, s));
            SUPR_HIP(hipStreamSynchronize(s));
            for (int i = 0; i < rows; ++i) y[i] = yf[i];
        } else {
            DevBuf dx = upload(std::vector<double>(x, x + cols), s), dy = device_vec<double>(rows);
            if (lanes == 0) M.lanes = rule_lanes(M, true);
            else if (lanes < 0) autotune_lanes(M, M.f64(), dx.as<double>(), dy.as<double>(), s);
            else M.lanes = lanes;
            launch_spmv(M, M.f64(), dx.as<double>(), dy.as<double>(), s);
            SUPR_HIP(hipMemcpyAsync(y, dy.as<void>(), sizeof(double) * rows, hipMemcpyDeviceToHost, s));
            SUPR_HIP(hipStreamSynchronize(s));
        }
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}

// Times the iteration kernels (see Solver::bench_kernels).  out: 8 doubles.
int supr_bench_kernels(const supr_lp* unscaled_form, const supr_lp* scaled_form,
                       const double* Dr, const double* Dc, const supr_pdlp_options* opt,
                       int32_t reps, double* out9) {
    using namespace supr;
    int rc = check_solve_args(unscaled_form, scaled_form, Dr, Dc, opt);
    if (rc != SUPR_OK) return rc;
    if (!out9 || reps < 1) { set_error("null output or reps < 1"); return SUPR_EINVAL; }
    try {
        SUPR_HIP(hipSetDevice(opt->device));
        supr_pdlp_options o = *opt;
        o.mode = SUPR_MODE_MIXED;
        ColumnRelabel R;   // the engine's layout, relabelled columns included
        const bool rl = relabel_columns(*unscaled_form, *scaled_form, Dc, opt->mode, R);
        Solver solver(rl ? R.U : *unscaled_form, rl ? R.S : *scaled_form, Dr, rl ? R.dc.data() : Dc, o);
        solver.setup();
        solver.bench_kernels(reps, out9);
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}

// Times a plain device copy of `bytes` (multiple of 16) reps times.
int supr_bench_copy(int64_t bytes, int32_t reps, int32_t device, double* ms_per_copy) {
    using namespace supr;
    if (bytes < 16 || bytes % 16 || reps < 1 || !ms_per_copy) {
        set_error("supr_bench_copy: bytes must be a positive multiple of 16, reps >= 1");
        return SUPR_EINVAL;
    }
    void *src = nullptr, *dst = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    try {
        SUPR_HIP(hipSetDevice(device));
        SUPR_HIP(hipMalloc(&src, bytes));
        SUPR_HIP(hipMalloc(&dst, bytes));
        SUPR_HIP(hipMemset(src, 0, bytes));
        SUPR_HIP(hipEventCreate(&e0));
        SUPR_HIP(hipEventCreate(&e1));
        const int64_t count = bytes / 16;
        // 1,024 blocks (4 per CU), grid-stride: the fastest of 1,024-8,192 on
        // MI355X, for 1 GiB each way (5.6 TB/s) and for 16-64 MB each way,
        // which stay in the 256 MB last-level cache (7.5-8.3 TB/s)
        const int grid = (int)std::min<int64_t>((count + kBlock - 1) / kBlock, 1024);
        hipLaunchKernelGGL(k_copy4, dim3(grid), dim3(kBlock), 0, nullptr, (const float4*)src, (float4*)dst, count);
        SUPR_HIP(hipEventRecord(e0, nullptr));
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_copy4, dim3(grid), dim3(kBlock), 0, nullptr, (const float4*)src, (float4*)dst, count);
        SUPR_HIP(hipEventRecord(e1, nullptr));
        SUPR_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SUPR_HIP(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_copy = ms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(src);
        (void)hipFree(dst);
        return SUPR_OK;
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (src) (void)hipFree(src);
        if (dst) (void)hipFree(dst);
        return map_exception();
    }
}

}  // extern "C"

void supr_release_cached_memory(void) { supr::release_cached(-1); }
