// prep.hip -- device preprocessing: stable CSR / CSC construction and the
// reference's scale() (src/scaling.cpp:25-84), plus the C ABI entry
// supr_scale_gpu.  Built with floating-point contraction off (the kernels'
// products and sums must round like the host's).
#include "prep.hpp"

#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "prep_kernels.hpp"

namespace supr {

namespace {

using namespace prep;

int grid_for(int64_t len, int per_block = kThreads) {
    const int64_t g = (len + per_block - 1) / per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, kGrid));
}

int lanes_for(int32_t rows, int32_t nnz) {
    const double avg = rows > 0 ? (double)nnz / rows : 0.0;
    return avg <= 2 ? 1 : avg <= 8 ? 4 : avg <= 32 ? 16 : 64;
}

#define PREP_LANES(L, ...)                                              \
    switch (L) {                                                        \
        case 1: { constexpr int LL = 1; __VA_ARGS__; } break;           \
        case 4: { constexpr int LL = 4; __VA_ARGS__; } break;           \
        case 16: { constexpr int LL = 16; __VA_ARGS__; } break;         \
        default: { constexpr int LL = 64; __VA_ARGS__; } break;         \
    }

// Exclusive scan of len int32 values (device) into out.
void exclusive_scan(const int32_t* in, int32_t* out, int64_t len, hipStream_t s) {
    if (len <= 0) return;
    const int64_t tiles = (len + kTile - 1) / kTile;
    DevBuf totals = device_vec<int32_t>(tiles);
    hipLaunchKernelGGL(kp_scan_tiles, dim3((unsigned)tiles), dim3(kThreads), 0, s, in, out, len,
                       totals.as<int32_t>());
    hipLaunchKernelGGL(kp_scan_totals, dim3(1), dim3(kThreads), 0, s, totals.as<int32_t>(),
                       (int32_t)tiles);
    hipLaunchKernelGGL(kp_scan_add, dim3(grid_for(len)), dim3(kThreads), 0, s, out, len,
                       totals.as<int32_t>());
    SUPR_HIP(hipGetLastError());
    SUPR_HIP(hipStreamSynchronize(s));   // totals is freed on return
}

// Pattern of the matrix whose entry k sits in row keys[k], column other[k]
// (device COO arrays; every key in [0, rows), checked by the caller).
void build_pattern(DevPattern& P, const int32_t* keys, const int32_t* other, int32_t nnz,
                   int32_t rows, int32_t cols, hipStream_t s) {
    P.rows = rows;
    P.cols = cols;
    P.nnz = nnz;
    P.ptr = device_vec<int32_t>((size_t)rows + 1);
    P.idx = device_vec<int32_t>(nnz);
    if (nnz == 0) {
        P.perm = device_vec<int32_t>(0);
        hipLaunchKernelGGL(kp_ptr_from_sorted, dim3(1), dim3(kThreads), 0, s, (const int32_t*)nullptr,
                           0, rows, P.ptr.as<int32_t>());
        SUPR_HIP(hipGetLastError());
    } else {
        int bits = 0;
        while (bits < 31 && (int64_t(1) << bits) < rows) ++bits;
        const int passes = std::max(1, (bits + kRadixBits - 1) / kRadixBits);
        const int32_t tiles = (int32_t)(((int64_t)nnz + kTile - 1) / kTile);
        DevBuf hist = device_vec<int32_t>((size_t)kRadix * tiles);
        DevBuf kbuf[2] = {device_vec<int32_t>(nnz), device_vec<int32_t>(nnz)};
        DevBuf vbuf[2] = {device_vec<int32_t>(nnz), device_vec<int32_t>(nnz)};
        const int32_t* kin = keys;
        const int32_t* vin = nullptr;
        int out = 0;
        for (int pass = 0; pass < passes; ++pass) {
            const int shift = pass * kRadixBits;
            hipLaunchKernelGGL(kp_radix_hist, dim3(tiles), dim3(kThreads), 0, s, kin, nnz, shift,
                               hist.as<int32_t>(), tiles);
            SUPR_HIP(hipGetLastError());
            exclusive_scan(hist.as<int32_t>(), hist.as<int32_t>(), (int64_t)kRadix * tiles, s);
            hipLaunchKernelGGL(kp_radix_scatter, dim3(tiles), dim3(kThreads), 0, s, kin, vin, nnz,
                               shift, (const int32_t*)hist.as<int32_t>(), tiles,
                               kbuf[out].as<int32_t>(), vbuf[out].as<int32_t>());
            SUPR_HIP(hipGetLastError());
            kin = kbuf[out].as<int32_t>();
            vin = vbuf[out].as<int32_t>();
            out ^= 1;
        }
        const int last = out ^ 1;
        hipLaunchKernelGGL(kp_ptr_from_sorted, dim3(grid_for((int64_t)nnz + 1)), dim3(kThreads), 0,
                           s, (const int32_t*)kbuf[last].as<int32_t>(), nnz, rows,
                           P.ptr.as<int32_t>());
        hipLaunchKernelGGL(kp_gather_idx, dim3(grid_for(nnz)), dim3(kThreads), 0, s,
                           (const int32_t*)vbuf[last].as<int32_t>(), nnz, other,
                           P.idx.as<int32_t>());
        SUPR_HIP(hipGetLastError());
        P.perm = std::move(vbuf[last]);
        SUPR_HIP(hipStreamSynchronize(s));   // the sort buffers are freed on return
    }
    // row-length range (ELL detection by the engine)
    std::vector<int32_t> hp((size_t)rows + 1);
    SUPR_HIP(hipMemcpyAsync(hp.data(), P.ptr.as<int32_t>(), sizeof(int32_t) * hp.size(),
                            hipMemcpyDeviceToHost, s));
    SUPR_HIP(hipStreamSynchronize(s));
    P.min_len = rows > 0 ? INT32_MAX : 0;
    P.max_len = 0;
    for (int32_t r = 0; r < rows; ++r) {
        const int32_t len = hp[(size_t)r + 1] - hp[r];
        P.min_len = std::min(P.min_len, len);
        P.max_len = std::max(P.max_len, len);
    }
}

DevBuf copy_of(const DevBuf& src, size_t bytes, hipStream_t s) {
    DevBuf d(bytes);
    if (bytes > 0)
        SUPR_HIP(hipMemcpyAsync(d.as<void>(), src.as<void>(), bytes, hipMemcpyDeviceToDevice, s));
    return d;
}

}  // namespace

void gather_values(const DevPattern& P, const double* coo_val, DevBuf& out, hipStream_t s) {
    out = device_vec<double>(P.nnz);
    if (P.nnz == 0) return;
    hipLaunchKernelGGL(kp_gather_val, dim3(grid_for(P.nnz)), dim3(kThreads), 0, s,
                       (const int32_t*)P.perm.as<int32_t>(), P.nnz, coo_val, out.as<double>());
    SUPR_HIP(hipGetLastError());
}

DevLp::DevLp(const supr_lp& lp, hipStream_t s, bool with_layout) : m(lp.m), n(lp.n), nnz(lp.nnz) {
    row = upload(lp.row_idx, nnz, s);
    col = upload(lp.col_idx, nnz, s);
    val = upload(lp.val, nnz, s);
    c = upload(lp.c, n, s);
    l = upload(lp.l, n, s);
    u = upload(lp.u, n, s);
    rL = upload(lp.rL, m, s);
    rU = upload(lp.rU, m, s);
    if (with_layout) layout(s);
}

DevLp::DevLp(int32_t m_, int32_t n_, int32_t nnz_, DevBuf&& row_, DevBuf&& col_, DevBuf&& val_,
             DevBuf&& c_, DevBuf&& l_, DevBuf&& u_, DevBuf&& rL_, DevBuf&& rU_, hipStream_t s)
    : m(m_), n(n_), nnz(nnz_), row(std::move(row_)), col(std::move(col_)), val(std::move(val_)),
      c(std::move(c_)), l(std::move(l_)), u(std::move(u_)), rL(std::move(rL_)), rU(std::move(rU_)) {
    layout(s);
}

void DevLp::layout(hipStream_t s) {
    build_pattern(A, row.as<int32_t>(), col.as<int32_t>(), nnz, m, n, s);
    build_pattern(AT, col.as<int32_t>(), row.as<int32_t>(), nnz, n, m, s);
    gather_values(A, val.as<double>(), vA, s);
    gather_values(AT, val.as<double>(), vAT, s);
}

void scale_on_device(const DevLp& lp, const char* integer, int ruiz_passes, DevScaled& out,
                     hipStream_t s) {
    const int32_t m = lp.m, n = lp.n, nnz = lp.nnz;
    const DevPattern& A = lp.A;
    const DevPattern& AT = lp.AT;
    DevBuf keep;
    if (integer) keep = upload(integer, n, s);
    const char* kp = integer ? keep.as<char>() : nullptr;
    out.vA = copy_of(lp.vA, sizeof(double) * nnz, s);
    out.vAT = copy_of(lp.vAT, sizeof(double) * nnz, s);
    out.Dr = device_vec<double>(m);
    out.Dc = device_vec<double>(n);
    hipLaunchKernelGGL(kp_fill, dim3(grid_for(m)), dim3(kThreads), 0, s, out.Dr.as<double>(), m, 1.0);
    hipLaunchKernelGGL(kp_fill, dim3(grid_for(n)), dim3(kThreads), 0, s, out.Dc.as<double>(), n, 1.0);
    DevBuf rs = device_vec<double>(m), cs = device_vec<double>(n);
    const int LA = lanes_for(m, nnz), LT = lanes_for(n, nnz);
    const int gA = grid_for(m, kThreads / LA), gT = grid_for(n, kThreads / LT);
    const int32_t* pA = A.ptr.as<int32_t>();
    const int32_t* iA = A.idx.as<int32_t>();
    const int32_t* pT = AT.ptr.as<int32_t>();
    const int32_t* iT = AT.idx.as<int32_t>();
    double* vA = out.vA.as<double>();
    double* vT = out.vAT.as<double>();
    // Ruiz (src/scaling.cpp:36-53): rows, then columns, ruiz_passes times
    for (int pass = 0; pass < ruiz_passes; ++pass) {
        const double* gprev = pass > 0 ? cs.as<double>() : nullptr;
        PREP_LANES(LA, hipLaunchKernelGGL((kp_ruiz<LL>), dim3(gA), dim3(kThreads), 0, s, m, pA, iA,
                                          vA, gprev, (const char*)nullptr, out.Dr.as<double>(),
                                          rs.as<double>()));
        PREP_LANES(LT, hipLaunchKernelGGL((kp_ruiz<LL>), dim3(gT), dim3(kThreads), 0, s, n, pT, iT,
                                          vT, (const double*)rs.as<double>(), kp,
                                          out.Dc.as<double>(), cs.as<double>()));
        SUPR_HIP(hipGetLastError());
    }
    // Pock-Chambolle (src/scaling.cpp:55-65)
    DevBuf pr = device_vec<double>(m), pc = device_vec<double>(n);
    const double* glast = ruiz_passes > 0 ? cs.as<double>() : nullptr;
    hipLaunchKernelGGL(kp_pc_short, dim3(grid_for(m)), dim3(kThreads), 0, s, m, pA, iA, vA, glast,
                       (const char*)nullptr, pr.as<double>());
    hipLaunchKernelGGL(kp_pc_long, dim3(grid_for(m, 4)), dim3(kThreads), 0, s, m, pA, iA, vA, glast,
                       (const char*)nullptr, pr.as<double>());
    hipLaunchKernelGGL(kp_pc_short, dim3(grid_for(n)), dim3(kThreads), 0, s, n, pT, iT, vT,
                       (const double*)nullptr, kp, pc.as<double>());
    hipLaunchKernelGGL(kp_pc_long, dim3(grid_for(n, 4)), dim3(kThreads), 0, s, n, pT, iT, vT,
                       (const double*)nullptr, kp, pc.as<double>());
    SUPR_HIP(hipGetLastError());
    PREP_LANES(LA, hipLaunchKernelGGL((kp_pc_apply<LL>), dim3(gA), dim3(kThreads), 0, s, m, pA, iA,
                                      vA, (const double*)pr.as<double>(),
                                      (const double*)pc.as<double>(), 1));
    PREP_LANES(LT, hipLaunchKernelGGL((kp_pc_apply<LL>), dim3(gT), dim3(kThreads), 0, s, n, pT, iT,
                                      vT, (const double*)pc.as<double>(),
                                      (const double*)pr.as<double>(), 0));
    hipLaunchKernelGGL(kp_mul, dim3(grid_for(m)), dim3(kThreads), 0, s, out.Dr.as<double>(),
                       (const double*)pr.as<double>(), m);
    hipLaunchKernelGGL(kp_mul, dim3(grid_for(n)), dim3(kThreads), 0, s, out.Dc.as<double>(),
                       (const double*)pc.as<double>(), n);
    SUPR_HIP(hipGetLastError());
    // scaled vectors (src/scaling.cpp:68-76)
    out.c = device_vec<double>(n); out.l = device_vec<double>(n); out.u = device_vec<double>(n);
    out.rL = device_vec<double>(m); out.rU = device_vec<double>(m);
    hipLaunchKernelGGL(kp_scale_cols, dim3(grid_for(n)), dim3(kThreads), 0, s, n,
                       (const double*)out.Dc.as<double>(), (const double*)lp.c.as<double>(),
                       (const double*)lp.l.as<double>(), (const double*)lp.u.as<double>(),
                       out.c.as<double>(), out.l.as<double>(), out.u.as<double>());
    hipLaunchKernelGGL(kp_scale_rows, dim3(grid_for(m)), dim3(kThreads), 0, s, m,
                       (const double*)out.Dr.as<double>(), (const double*)lp.rL.as<double>(),
                       (const double*)lp.rU.as<double>(), out.rL.as<double>(), out.rU.as<double>());
    SUPR_HIP(hipGetLastError());
    SUPR_HIP(hipStreamSynchronize(s));   // temporaries are freed on return
}

void build_form_on_device(const DevLp& lp, DevForm& f, hipStream_t s) {
    const int32_t m = lp.m, n = lp.n, nnz = lp.nnz;
    f.m_orig = m;
    f.n_orig = n;
    DevBuf fE = device_vec<int32_t>(m), fI = device_vec<int32_t>(m), fR = device_vec<int32_t>(m);
    DevBuf sE = device_vec<int32_t>(m), sI = device_vec<int32_t>(m), sR = device_vec<int32_t>(m);
    int32_t tot[3] = {0, 0, 0};
    if (m > 0) {
        hipLaunchKernelGGL(kf_flags, dim3(grid_for(m)), dim3(kThreads), 0, s, m,
                           lp.rL.as<const double>(), lp.rU.as<const double>(), fE.as<int32_t>(),
                           fI.as<int32_t>(), fR.as<int32_t>());
        SUPR_HIP(hipGetLastError());
        exclusive_scan(fE.as<int32_t>(), sE.as<int32_t>(), m, s);
        exclusive_scan(fI.as<int32_t>(), sI.as<int32_t>(), m, s);
        exclusive_scan(fR.as<int32_t>(), sR.as<int32_t>(), m, s);
        DevBuf dt = device_vec<int32_t>(3);
        hipLaunchKernelGGL(kf_totals, dim3(1), dim3(64), 0, s, m, sE.as<const int32_t>(),
                           fE.as<const int32_t>(), sI.as<const int32_t>(), fI.as<const int32_t>(),
                           sR.as<const int32_t>(), fR.as<const int32_t>(), dt.as<int32_t>());
        SUPR_HIP(hipGetLastError());
        SUPR_HIP(hipMemcpyAsync(tot, dt.as<void>(), sizeof(tot), hipMemcpyDeviceToHost, s));
        SUPR_HIP(hipStreamSynchronize(s));
    }
    const int32_t n_eq = tot[0], n_ineq = tot[1];
    f.n_slack = tot[2];
    f.m2 = n_eq + n_ineq;
    f.n2 = n + f.n_slack;
    f.row_perm = device_vec<int32_t>(f.m2);
    f.row_signs = device_vec<double>(f.m2);
    f.rL = device_vec<double>(f.m2);
    f.rU = device_vec<double>(f.m2);
    f.c = device_vec<double>(f.n2);
    f.l = device_vec<double>(f.n2);
    f.u = device_vec<double>(f.n2);
    if (n > 0) {
        SUPR_HIP(hipMemcpyAsync(f.c.as<void>(), lp.c.as<void>(), sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        SUPR_HIP(hipMemcpyAsync(f.l.as<void>(), lp.l.as<void>(), sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        SUPR_HIP(hipMemcpyAsync(f.u.as<void>(), lp.u.as<void>(), sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    }
    if (f.n_slack > 0)
        hipLaunchKernelGGL(kp_fill, dim3(grid_for(f.n_slack)), dim3(kThreads), 0, s,
                           f.c.as<double>() + n, f.n_slack, 0.0);
    DevBuf pos_of = device_vec<int32_t>(m), slack_pos = device_vec<int32_t>(f.n_slack);
    if (m > 0)
        hipLaunchKernelGGL(kf_rows, dim3(grid_for(m)), dim3(kThreads), 0, s, m, n, n_eq,
                           lp.rL.as<const double>(), lp.rU.as<const double>(),
                           sE.as<const int32_t>(), sI.as<const int32_t>(), sR.as<const int32_t>(),
                           f.row_perm.as<int32_t>(), f.row_signs.as<double>(), f.rL.as<double>(),
                           f.rU.as<double>(), pos_of.as<int32_t>(), slack_pos.as<int32_t>(),
                           f.l.as<double>(), f.u.as<double>());
    SUPR_HIP(hipGetLastError());
    // entries: the kept ones in COO order, then one per slack
    int32_t kept = 0;
    DevBuf fK = device_vec<int32_t>(nnz), q = device_vec<int32_t>(nnz);
    if (nnz > 0) {
        hipLaunchKernelGGL(kf_keep, dim3(grid_for(nnz)), dim3(kThreads), 0, s, nnz,
                           lp.row.as<const int32_t>(), pos_of.as<const int32_t>(), fK.as<int32_t>());
        SUPR_HIP(hipGetLastError());
        exclusive_scan(fK.as<int32_t>(), q.as<int32_t>(), nnz, s);
        int32_t last[2];
        SUPR_HIP(hipMemcpyAsync(&last[0], q.as<int32_t>() + nnz - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        SUPR_HIP(hipMemcpyAsync(&last[1], fK.as<int32_t>() + nnz - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        SUPR_HIP(hipStreamSynchronize(s));
        kept = last[0] + last[1];
    }
    f.nnz2 = kept + f.n_slack;
    f.row = device_vec<int32_t>(f.nnz2);
    f.col = device_vec<int32_t>(f.nnz2);
    f.val = device_vec<double>(f.nnz2);
    if (nnz > 0)
        hipLaunchKernelGGL(kf_entries, dim3(grid_for(nnz)), dim3(kThreads), 0, s, nnz,
                           lp.row.as<const int32_t>(), lp.col.as<const int32_t>(),
                           lp.val.as<const double>(), pos_of.as<const int32_t>(),
                           f.row_signs.as<const double>(), fK.as<const int32_t>(),
                           q.as<const int32_t>(), f.row.as<int32_t>(), f.col.as<int32_t>(),
                           f.val.as<double>());
    if (f.n_slack > 0)
        hipLaunchKernelGGL(kf_slack_entries, dim3(grid_for(f.n_slack)), dim3(kThreads), 0, s,
                           f.n_slack, kept, n, slack_pos.as<const int32_t>(), f.row.as<int32_t>(),
                           f.col.as<int32_t>(), f.val.as<double>());
    SUPR_HIP(hipGetLastError());
    SUPR_HIP(hipStreamSynchronize(s));   // temporaries are freed on return
}

void verify_products_on_device(const DevLp& lp, const double* x, const double* y, double* ax,
                               double* aty, hipStream_t s) {
    auto run = [&](const DevPattern& P, const DevBuf& v, const double* in, double* out) {
        if (P.rows == 0) return;
        hipLaunchKernelGGL(kp_spmv_seq_short, dim3(grid_for(P.rows)), dim3(kThreads), 0, s, P.rows,
                           P.ptr.as<const int32_t>(), P.idx.as<const int32_t>(), v.as<const double>(),
                           in, out);
        if (P.max_len > kLongRow)
            hipLaunchKernelGGL(kp_spmv_seq_long, dim3(grid_for(P.rows, kThreads / 64)), dim3(kThreads),
                               0, s, P.rows, P.ptr.as<const int32_t>(), P.idx.as<const int32_t>(),
                               v.as<const double>(), in, out);
        SUPR_HIP(hipGetLastError());
    };
    run(lp.A, lp.vA, x, ax);
    run(lp.AT, lp.vAT, y, aty);
}

double device_absmax(const double* v, int32_t len, hipStream_t s) {
    if (len <= 0) return 0.0;
    const int g = grid_for(len);
    DevBuf part = device_vec<double>(g);
    hipLaunchKernelGGL(kp_absmax, dim3(g), dim3(kThreads), 0, s, v, len, part.as<double>());
    SUPR_HIP(hipGetLastError());
    std::vector<double> h(g);
    SUPR_HIP(hipMemcpyAsync(h.data(), part.as<void>(), sizeof(double) * g, hipMemcpyDeviceToHost, s));
    SUPR_HIP(hipStreamSynchronize(s));
    double mx = 0.0;
    for (double x : h) mx = mx < x ? x : mx;
    return mx;
}

void scatter_to_coo(const DevPattern& P, const double* v, double* out, hipStream_t s) {
    if (P.nnz == 0) return;
    hipLaunchKernelGGL(kp_scatter_coo, dim3(grid_for(P.nnz)), dim3(kThreads), 0, s,
                       (const int32_t*)P.perm.as<int32_t>(), P.nnz, v, out);
    SUPR_HIP(hipGetLastError());
}

}  // namespace supr

extern "C" {

int supr_scale_gpu(const supr_lp* lp, const char* integer, int32_t ruiz_passes, int32_t device,
                   double* val_s, double* c_s, double* rL_s, double* rU_s, double* l_s,
                   double* u_s, double* Dr, double* Dc, double* seconds3) {
    using namespace supr;
    if (!check_lp(lp, "supr_scale_gpu")) return SUPR_EINVAL;
    if (ruiz_passes < 0) { set_error("supr_scale_gpu: negative ruiz_passes"); return SUPR_EINVAL; }
    if ((lp->nnz > 0 && !val_s) || (lp->n > 0 && (!c_s || !l_s || !u_s || !Dc)) ||
        (lp->m > 0 && (!rL_s || !rU_s || !Dr))) {
        set_error("supr_scale_gpu: null output");
        return SUPR_EINVAL;
    }
    hipStream_t s = nullptr;
    try {
        SUPR_HIP(hipSetDevice(device));
        SUPR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamOwner {
            hipStream_t s;
            ~StreamOwner() { (void)hipStreamDestroy(s); }
        } own{s};
        const auto t0 = Clock::now();
        DevLp D(*lp, s);
        const double t_layout = seconds_since(t0);
        const auto t1 = Clock::now();
        DevScaled S;
        scale_on_device(D, integer, ruiz_passes, S, s);
        const double t_scale = seconds_since(t1);
        const auto t2 = Clock::now();
        const int32_t m = lp->m, n = lp->n, nnz = lp->nnz;
        if (nnz > 0) {
            DevBuf coo = device_vec<double>(nnz);
            scatter_to_coo(D.A, S.vA.as<double>(), coo.as<double>(), s);
            SUPR_HIP(hipMemcpyAsync(val_s, coo.as<void>(), sizeof(double) * nnz, hipMemcpyDeviceToHost, s));
            SUPR_HIP(hipStreamSynchronize(s));
        }
        auto down = [&](double* dst, const DevBuf& src, int32_t len) {
            if (len > 0)
                SUPR_HIP(hipMemcpyAsync(dst, src.as<void>(), sizeof(double) * len, hipMemcpyDeviceToHost, s));
        };
        down(c_s, S.c, n); down(l_s, S.l, n); down(u_s, S.u, n); down(Dc, S.Dc, n);
        down(rL_s, S.rL, m); down(rU_s, S.rU, m); down(Dr, S.Dr, m);
        SUPR_HIP(hipStreamSynchronize(s));
        if (seconds3) {
            seconds3[0] = t_layout;
            seconds3[1] = t_scale;
            seconds3[2] = seconds_since(t2);
        }
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}

// build_pdlp_form on the device; outputs as supr_pdlp_form's.
int supr_pdlp_form_gpu(const supr_lp* lp, int32_t device, int32_t* row_idx2, int32_t* col_idx2,
                       double* val2, double* c2, double* rL2, double* rU2, double* l2, double* u2,
                       int32_t* row_perm, double* row_signs, int32_t* dims4) {
    using namespace supr;
    if (!check_lp(lp, "supr_pdlp_form_gpu")) return SUPR_EINVAL;
    hipStream_t s = nullptr;
    try {
        SUPR_HIP(hipSetDevice(device));
        SUPR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamOwner {
            hipStream_t s;
            ~StreamOwner() { (void)hipStreamDestroy(s); }
        } own{s};
        DevLp D(*lp, s, false);
        DevForm F;
        build_form_on_device(D, F, s);
        if (dims4) { dims4[0] = F.m2; dims4[1] = F.n2; dims4[2] = F.nnz2; dims4[3] = F.n_slack; }
        auto down = [&](void* dst, const DevBuf& src, size_t bytes) {
            if (bytes > 0) {
                if (!dst) throw BadInput("supr_pdlp_form_gpu: null output");
                SUPR_HIP(hipMemcpyAsync(dst, src.as<void>(), bytes, hipMemcpyDeviceToHost, s));
            }
        };
        down(row_idx2, F.row, sizeof(int32_t) * F.nnz2);
        down(col_idx2, F.col, sizeof(int32_t) * F.nnz2);
        down(val2, F.val, sizeof(double) * F.nnz2);
        down(c2, F.c, sizeof(double) * F.n2);
        down(l2, F.l, sizeof(double) * F.n2);
        down(u2, F.u, sizeof(double) * F.n2);
        down(rL2, F.rL, sizeof(double) * F.m2);
        down(rU2, F.rU, sizeof(double) * F.m2);
        down(row_perm, F.row_perm, sizeof(int32_t) * F.m2);
        down(row_signs, F.row_signs, sizeof(double) * F.m2);
        SUPR_HIP(hipStreamSynchronize(s));
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}

// verify() with its two products on the device (left-to-right row sums, as
// the host's) and the O(m + n) tail on the host: the same report, bit for bit.
int supr_verify_gpu(const supr_lp* lp, const double* x, const double* y, int32_t device,
                    double* out7) {
    using namespace supr;
    if (!check_lp(lp, "supr_verify_gpu")) return SUPR_EINVAL;
    if ((lp->n > 0 && !x) || (lp->m > 0 && !y) || !out7) {
        set_error("supr_verify_gpu: null x, y or out");
        return SUPR_EINVAL;
    }
    hipStream_t s = nullptr;
    try {
        SUPR_HIP(hipSetDevice(device));
        SUPR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamOwner {
            hipStream_t s;
            ~StreamOwner() { (void)hipStreamDestroy(s); }
        } own{s};
        DevLp D(*lp, s);
        DevBuf dx = upload(x, lp->n, s), dy = upload(y, lp->m, s);
        DevBuf ax = device_vec<double>(lp->m), aty = device_vec<double>(lp->n);
        verify_products_on_device(D, dx.as<double>(), dy.as<double>(), ax.as<double>(), aty.as<double>(), s);
        std::vector<double> hax(lp->m), haty(lp->n);
        if (lp->m > 0)
            SUPR_HIP(hipMemcpyAsync(hax.data(), ax.as<void>(), sizeof(double) * lp->m, hipMemcpyDeviceToHost, s));
        if (lp->n > 0)
            SUPR_HIP(hipMemcpyAsync(haty.data(), aty.as<void>(), sizeof(double) * lp->n, hipMemcpyDeviceToHost, s));
        SUPR_HIP(hipStreamSynchronize(s));
        verify_products(*lp, x, y, hax.data(), haty.data(), out7);
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}

// CSR of A (transposed = 0) or of A^T (transposed = 1) built on the device.
int supr_csr_gpu(const supr_lp* lp, int32_t transposed, int32_t device, int32_t* ptr,
                 int32_t* idx, int32_t* perm) {
    using namespace supr;
    if (!check_lp(lp, "supr_csr_gpu")) return SUPR_EINVAL;
    const int32_t rows = transposed ? lp->n : lp->m;
    if (!ptr || (lp->nnz > 0 && (!idx || !perm))) {
        set_error("supr_csr_gpu: null output");
        return SUPR_EINVAL;
    }
    hipStream_t s = nullptr;
    try {
        SUPR_HIP(hipSetDevice(device));
        SUPR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamOwner {
            hipStream_t s;
            ~StreamOwner() { (void)hipStreamDestroy(s); }
        } own{s};
        DevBuf r = upload(lp->row_idx, lp->nnz, s), c = upload(lp->col_idx, lp->nnz, s);
        DevPattern P;
        build_pattern(P, (transposed ? c : r).as<int32_t>(), (transposed ? r : c).as<int32_t>(),
                      lp->nnz, rows, transposed ? lp->m : lp->n, s);
        SUPR_HIP(hipMemcpyAsync(ptr, P.ptr.as<void>(), sizeof(int32_t) * ((size_t)rows + 1),
                                hipMemcpyDeviceToHost, s));
        if (lp->nnz > 0) {
            SUPR_HIP(hipMemcpyAsync(idx, P.idx.as<void>(), sizeof(int32_t) * lp->nnz, hipMemcpyDeviceToHost, s));
            SUPR_HIP(hipMemcpyAsync(perm, P.perm.as<void>(), sizeof(int32_t) * lp->nnz, hipMemcpyDeviceToHost, s));
        }
        SUPR_HIP(hipStreamSynchronize(s));
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}

}  // extern "C"
