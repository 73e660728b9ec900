// pdlp_kernels.hpp -- the single-LP engine's kernels (plain and device-
// resident); included by pdlp_engine.hip only.  Helpers, layouts and the
// description of the SpMV schedules are in pdlp_device.hpp.
#pragma once
#include "pdlp_device.hpp"

namespace supr {

// ------------------------------------------------------- plain kernels --

// y = M x  (the reference's sov_spmv_csr_* pair, generalised to L lanes)
template <int L, typename V, typename X, typename Y>
__global__ __launch_bounds__(kBlock) void k_spmv(MatView<V> M, const X* __restrict__ x,
                                                 Y* __restrict__ y) {
    if (staged_rows<L>(M.rows)) {
        for_row_tiles<L, 1, V>(
            M.rows, xcd_block(), gridDim.x,
            [&](int row, int lane, bool valid, V (&v)[1]) { v[0] = row_dot<L, V>(M, x, row, lane, valid); },
            [&](int row, const V (&v)[1]) { y[row] = (Y)v[0]; });
        return;
    }
    for_rows<L>(M.rows, [&](int row, int lane, bool valid) {
        const V s = row_dot<L, V>(M, x, row, lane, valid);
        if (valid && lane == 0) y[row] = (Y)s;
    });
}

// xn = clip(x - tau (c - aty), l, u)          gpu/pdlp_gpu_solve.cu:71-80
__global__ __launch_bounds__(kBlock) void k_plain_primal(int n, const double* __restrict__ x,
                                                         const double* __restrict__ c,
                                                         const double* __restrict__ aty,
                                                         const double* __restrict__ l,
                                                         const double* __restrict__ u,
                                                         double tau, double* __restrict__ xn) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j < n) xn[j] = smin(smax(x[j] - tau * (c[j] - aty[j]), l[j]), u[j]);
}

// yn = y + sigma (b - (2 axn - ax)), clipped at 0 on inequality rows   :82-92
__global__ __launch_bounds__(kBlock) void k_plain_dual(int m, const double* __restrict__ y,
                                                       const double* __restrict__ b,
                                                       const double* __restrict__ axn,
                                                       const double* __restrict__ ax,
                                                       const uint8_t* __restrict__ ineq,
                                                       double sigma, double* __restrict__ yn) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < m) {
        const double v = y[i] + sigma * (b[i] - (2.0 * axn[i] - ax[i]));
        yn[i] = (ineq[i] && v < 0.0) ? 0.0 : v;
    }
}

// Per-block partials of sum (xn-x)^2 and of sum (yn-y)^2, sum (yn-y)(axn-ax):
// part[b], part[kMaxPlain + b], part[2 kMaxPlain + b]          :105-131
constexpr int kMaxPlain = 256;
__global__ __launch_bounds__(kBlock) void k_plain_stats_n(int n, const double* __restrict__ x,
                                                          const double* __restrict__ xn,
                                                          double* __restrict__ part) {
    double v[1] = {0.0};
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const double d = xn[j] - x[j];
        v[0] += d * d;
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ __launch_bounds__(kBlock) void k_plain_stats_m(int m, const double* __restrict__ y,
                                                          const double* __restrict__ yn,
                                                          const double* __restrict__ ax,
                                                          const double* __restrict__ axn,
                                                          double* __restrict__ part) {
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        const double d = yn[i] - y[i];
        v[0] += d * d;
        v[1] += d * (axn[i] - ax[i]);
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        part[kMaxPlain + blockIdx.x] = v[0];
        part[2 * kMaxPlain + blockIdx.x] = v[1];
    }
}

// y += a x                                                     :133-138
__global__ __launch_bounds__(kBlock) void k_plain_axpy(int len, double a,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ y) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j < len) y[j] += a * x[j];
}

// Plain stream copy (16-byte vectors, grid-stride): the reference point for
// achievable bandwidth and for calibrating HBM traffic counters.
__global__ __launch_bounds__(kBlock) void k_copy4(const float4* __restrict__ src,
                                                  float4* __restrict__ dst, int64_t count) {
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < count;
         k += (int64_t)gridDim.x * kBlock)
        dst[k] = src[k];
}

// ------------------------------------------- device-resident PDHG loop --
//
// Launch sequence of one graph replay (G attempts, G even):
//   k_primal(slot 0), k_dual, k_primal(slot 1), k_dual, ... , then the check
//   group k_check_prep, k_check_kkt, k_kkt_fin, k_check_restart, k_restart_fin.
// The step decision of attempt t is not a launch of its own: every block of
// the next kernel (k_primal of attempt t+1, or k_check_prep) recomputes it
// from the dual partials of attempt t and the primal sum k_dual's block 0
// stored -- same data, same order, same bits in every block.  To let blocks
// read the old state while block 0 publishes the new one, the state is
// double-buffered: k_primal of graph slot k reads S[k & 1] and writes
// S[(k + 1) & 1]; k_dual reads S[(k + 1) & 1]; the check group works on S[1]
// and k_restart_fin copies S[1] back to S[0] for the next replay.

// Primal half of one attempt, over the columns (rows of A^T):
//   aty = A^T y (formed every attempt, not stored: a rejected attempt forms
//   the same product of the same y again, which costs less than writing it
//   on every accepted one), [after an accept] xsum += eta_acc x (pending
//   average term), xn = clip(x - tau (c - aty), l, u); partial sums of
//   (xn - x)^2 in FP64.
// Pre: a one-pass launch (one_pass_rows); the lane's first matrix slots and
// its row's epilogue operands (both x buffers, c, xsum, l, u) are loaded
// before the decision prologue.  The same arithmetic, so the same bits.
template <int L, typename T, bool Pre>
__global__ __launch_bounds__(kBlock) void k_primal(DevState* __restrict__ S, int p,
                                                   MatView<T> AT, IterBufs<T> B,
                                                   const double* __restrict__ part_m, int nb_m,
                                                   const double* __restrict__ dx2_slot,
                                                   double* __restrict__ part_n) {
    __shared__ DevState Q;
    int prow = 0, plane = 0;
    bool pvalid = false;
    PreSlots<T> P;
    T px[2] = {T(0), T(0)}, pc = T(0), ps = T(0), pl = T(0), pu = T(0);
    if constexpr (Pre) {
        prow = xcd_block() * (kBlock / L) + (int)threadIdx.x / L;
        plane = (int)threadIdx.x % L;
        pvalid = prow < AT.rows;
        pre_slots<L>(AT, prow, plane, pvalid, P);
        if (pvalid && plane == 0) {
            px[0] = B.x[0][prow];
            px[1] = B.x[1][prow];
            pc = B.c[prow];
            ps = B.xsum[prow];
            if (B.l) { pl = B.l[prow]; pu = B.u[prow]; }
        }
    }
    // the publisher is the last block: its slot has the fewest rows, so the
    // step powers it forms before its rows stay off the launch's critical
    // path.  (Formed after the rows instead, the publisher flag and S stay
    // live through the row loops, and the FP64 variant then reloads kernel
    // arguments inside its pass loop: 40% slower on the power-law LP.)
    const bool publisher = blockIdx.x == gridDim.x - 1;
    load_and_decide(S, p, part_m, nb_m, dx2_slot, Q, publisher);
    if (!Q.active) return;
    if (publisher && threadIdx.x == 0) {
        S[p ^ 1].decide_pending = 1;   // k_dual will run
        publish_step_powers(S, p, Q);
    }
    const int cur = Q.cur;
    const bool pending = Q.pending != 0;
    const T eta_acc = (T)Q.eta_acc;
    const T tau = (T)(Q.eta / Q.omega);
    const T* __restrict__ x = B.x[cur];
    T* __restrict__ xn = B.x[cur ^ 1];
    const T* __restrict__ y = B.y[cur];
    double acc[1] = {0.0};
    auto epilogue = [&](int row, T a) {
        const T xj = x[row];
        if (pending) B.xsum[row] += eta_acc * xj;
        const T z = xj - tau * (B.c[row] - a);
        // B.l null: l = +0 and u = +inf everywhere, and smin(smax(z, +0), +inf)
        // is smax(z, +0) bit for bit -- 2 of the 7 per-column streams skipped
        const T v = B.l ? smin(smax(z, B.l[row]), B.u[row]) : smax(z, T(0));
        xn[row] = v;
        if constexpr (sizeof(T) == 4)
            reinterpret_cast<float2*>(B.xd)[row] = make_float2(v, v - xj);
        const double d = (double)v - (double)xj;
        acc[0] += d * d;
    };
    if constexpr (Pre) {
        const T a = row_dot_pre<L, false, T>(AT, y, prow, P);
        if (pvalid && plane == 0) {
            // the epilogue above with the operands loaded before the decision
            const T xj = px[cur];
            if (pending) B.xsum[prow] = ps + eta_acc * xj;
            const T z = xj - tau * (pc - a);
            const T v = B.l ? smin(smax(z, pl), pu) : smax(z, T(0));
            xn[prow] = v;
            if constexpr (sizeof(T) == 4)
                reinterpret_cast<float2*>(B.xd)[prow] = make_float2(v, v - xj);
            const double d = (double)v - (double)xj;
            acc[0] += d * d;
        }
    } else if (L == 1 && AT.pairs) {
        // two adjacent rows per lane: the epilogue above on both, its streams
        // read and written as one 8- or 16-byte access per lane
        using V2 = typename Pair<T>::type;
        for (int q = xcd_block() * kBlock + (int)threadIdx.x; 2 * q < AT.rows; q += (int)gridDim.x * kBlock) {
            const int r = 2 * q;
            T a[2];
            ell_pair_dot(AT, y, r, a);
            const V2 xv = *reinterpret_cast<const V2*>(x + r);
            const T xj[2] = {xv.x, xv.y};
            if (pending) {
                V2 sv = *reinterpret_cast<const V2*>(B.xsum + r);
                sv.x += eta_acc * xj[0];
                sv.y += eta_acc * xj[1];
                *reinterpret_cast<V2*>(B.xsum + r) = sv;
            }
            const V2 cv = *reinterpret_cast<const V2*>(B.c + r);
            const T z[2] = {xj[0] - tau * (cv.x - a[0]), xj[1] - tau * (cv.y - a[1])};
            T v[2];
            if (B.l) {
                const V2 lv = *reinterpret_cast<const V2*>(B.l + r);
                const V2 uv = *reinterpret_cast<const V2*>(B.u + r);
                v[0] = smin(smax(z[0], lv.x), uv.x);
                v[1] = smin(smax(z[1], lv.y), uv.y);
            } else {
                v[0] = smax(z[0], T(0));
                v[1] = smax(z[1], T(0));
            }
            V2 out;
            out.x = v[0];
            out.y = v[1];
            *reinterpret_cast<V2*>(xn + r) = out;
            if constexpr (sizeof(T) == 4)
                reinterpret_cast<float4*>(B.xd)[q] = make_float4(v[0], v[0] - xj[0], v[1], v[1] - xj[1]);
            // an explicit fma: written acc += d * d, the two squares' sum is
            // contracted by the compiler's choice of which multiply to fuse,
            // and that choice differed between kernels with the same source
            // (k_primal_pairs fused the first square, this loop the second)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const double d = (double)v[h] - (double)xj[h];
                acc[0] = fma(d, d, acc[0]);
            }
        }
    } else if (staged_rows<L>(AT.rows)) {
        for_row_tiles<L, 1, T>(
            AT.rows, xcd_block(), gridDim.x,
            [&](int row, int lane, bool valid, T (&v)[1]) { v[0] = row_dot<L, T>(AT, y, row, lane, valid); },
            [&](int row, const T (&v)[1]) { epilogue(row, v[0]); });
    } else {
        for_rows<L>(AT.rows, [&](int row, int lane, bool valid) {
            const T a = row_dot<L, T>(AT, y, row, lane, valid);
            if (valid && lane == 0) epilogue(row, a);
        });
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) part_n[blockIdx.x] = acc[0];
}

// The primal half over an ELL A^T with one lane per row in one pass of row
// pairs (MatView::pairs and rows <= 2 x grid x kBlock; the 1M-variable
// transportation LP's A^T): lane q takes rows 2q and 2q + 1, and loads the
// pair's first two index slots before the decision prologue, so the gathers
// of y start as soon as the decision is known instead of one memory latency
// later; the rest of the pair's operands are loaded after the decision,
// beside the gathers.  The terms, epilogue and partial sums are those of
// k_primal's row-pair path (one pair per lane), so the same bits.  A kernel
// of its own: only this path sets its register count (64 VGPRs in FP64; the
// FP32 loop keeps k_primal's path, where the early slots took it to 66, a
// wavefront per SIMD less).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_primal_pairs(DevState* __restrict__ S, int p,
                                                         MatView<T> AT, IterBufs<T> B,
                                                         const double* __restrict__ part_m, int nb_m,
                                                         const double* __restrict__ dx2_slot,
                                                         double* __restrict__ part_n) {
    using V2 = typename Pair<T>::type;
    __shared__ DevState Q;
    const size_t R = (size_t)AT.rows;
    const int q = xcd_block() * kBlock + (int)threadIdx.x;
    const int r = 2 * q;
    const bool valid = r < AT.rows;
    int c0[2] = {0, 0}, c1[2] = {0, 0};   // slots 0 and 1 of rows r, r + 1
    auto slot = [&](int k, int (&c)[2]) {
        if (AT.idx16) {
            const ushort2 u = *reinterpret_cast<const ushort2*>(AT.idx16 + (size_t)k * R + r);
            c[0] = u.x;
            c[1] = u.y;
        } else {
            const int2 u = *reinterpret_cast<const int2*>(AT.idx + (size_t)k * R + r);
            c[0] = u.x;
            c[1] = u.y;
        }
    };
    if (valid && AT.ell_w > 0) slot(0, c0);
    if (valid && AT.ell_w > 1) slot(1, c1);
    const bool publisher = blockIdx.x == gridDim.x - 1;
    load_and_decide(S, p, part_m, nb_m, dx2_slot, Q, publisher);
    if (!Q.active) return;
    if (publisher && threadIdx.x == 0) {
        S[p ^ 1].decide_pending = 1;   // k_dual will run
        publish_step_powers(S, p, Q);
    }
    const int cur = Q.cur;
    const bool pending = Q.pending != 0;
    const T eta_acc = (T)Q.eta_acc;
    const T tau = (T)(Q.eta / Q.omega);
    const T* __restrict__ x = B.x[cur];
    T* __restrict__ xn = B.x[cur ^ 1];
    const T* __restrict__ y = B.y[cur];
    double acc[1] = {0.0};
    if (valid) {
        // ell_pair_dot's terms in its order: slot 0, slot 1, then the rest
        T a[2] = {T(0), T(0)};
        bool rebuilt = false;
        if constexpr (std::is_same<T, double>::value) rebuilt = AT.csc != nullptr;
        if (rebuilt) {
            if constexpr (std::is_same<T, double>::value) {
                const double2 f = *reinterpret_cast<const double2*>(AT.rsc + r);
                if (AT.ell_w > 0) {
                    a[0] += (AT.csc[c0[0]] * f.x) * y[c0[0]];
                    a[1] += (AT.csc[c0[1]] * f.y) * y[c0[1]];
                }
                if (AT.ell_w > 1) {
                    a[0] += (AT.csc[c1[0]] * f.x) * y[c1[0]];
                    a[1] += (AT.csc[c1[1]] * f.y) * y[c1[1]];
                }
                for (int k = 2; k < AT.ell_w; ++k) {
                    int c[2];
                    slot(k, c);
                    a[0] += (AT.csc[c[0]] * f.x) * y[c[0]];
                    a[1] += (AT.csc[c[1]] * f.y) * y[c[1]];
                }
            }
        } else {
            auto vals = [&](int k) { return *reinterpret_cast<const V2*>(AT.val + (size_t)k * R + r); };
            if (AT.ell_w > 0) {
                const V2 v = vals(0);
                a[0] += v.x * y[c0[0]];
                a[1] += v.y * y[c0[1]];
            }
            if (AT.ell_w > 1) {
                const V2 v = vals(1);
                a[0] += v.x * y[c1[0]];
                a[1] += v.y * y[c1[1]];
            }
            for (int k = 2; k < AT.ell_w; ++k) {
                int c[2];
                slot(k, c);
                const V2 v = vals(k);
                a[0] += v.x * y[c[0]];
                a[1] += v.y * y[c[1]];
            }
        }
        // k_primal's row-pair epilogue
        const V2 xv = *reinterpret_cast<const V2*>(x + r);
        const T xj[2] = {xv.x, xv.y};
        if (pending) {
            V2 sv = *reinterpret_cast<const V2*>(B.xsum + r);
            sv.x += eta_acc * xj[0];
            sv.y += eta_acc * xj[1];
            *reinterpret_cast<V2*>(B.xsum + r) = sv;
        }
        const V2 cv = *reinterpret_cast<const V2*>(B.c + r);
        const T z[2] = {xj[0] - tau * (cv.x - a[0]), xj[1] - tau * (cv.y - a[1])};
        T v[2];
        if (B.l) {
            const V2 lv = *reinterpret_cast<const V2*>(B.l + r);
            const V2 uv = *reinterpret_cast<const V2*>(B.u + r);
            v[0] = smin(smax(z[0], lv.x), uv.x);
            v[1] = smin(smax(z[1], lv.y), uv.y);
        } else {
            v[0] = smax(z[0], T(0));
            v[1] = smax(z[1], T(0));
        }
        V2 out;
        out.x = v[0];
        out.y = v[1];
        *reinterpret_cast<V2*>(xn + r) = out;
        if constexpr (sizeof(T) == 4)
            reinterpret_cast<float4*>(B.xd)[q] = make_float4(v[0], v[0] - xj[0], v[1], v[1] - xj[1]);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const double d = (double)v[h] - (double)xj[h];
            acc[0] = fma(d, d, acc[0]);   // see k_primal's row-pair path
        }
    }
    block_sum<1>(acc);
    if (threadIdx.x == 0) part_n[blockIdx.x] = acc[0];
}

// Dual half of one attempt, over the rows of A:
//   axn = A xn, ad = A (xn - x);  [pending] ysum += eta_acc y;
//   yn = proj(y + sigma (b - (axn + ad)));  partials of (yn-y)^2 and
//   (yn-y)'ad.  The last block also sums k_primal's partials into *dx2_slot.
// Pre: 1, a one-pass launch; the lane's first matrix slots and its row's
// epilogue operands (both y and A x buffers, b, ysum, the row kind) are
// loaded before the state.  2 (CSR): the same with the lane's first chunk
// of entries (row_dot_flat, row_dot_diff_flat) instead of the slots.  The
// same arithmetic, so the same bits.
template <int L, typename T, int Pre>
__global__ __launch_bounds__(kBlock) void k_dual(const DevState* __restrict__ S, int p1,
                                                 MatView<T> A, IterBufs<T> B,
                                                 const double* __restrict__ part_n, int nb_n,
                                                 double* __restrict__ dx2_slot,
                                                 double* __restrict__ part_m) {
    constexpr bool kDiff = sizeof(T) == 4;
    int prow = 0, plane = 0;
    bool pvalid = false;
    PreSlots<T> P;
    T py[2] = {T(0), T(0)}, pax[2] = {T(0), T(0)}, pb = T(0), ps = T(0);
    uint8_t pineq = 0;
    FlatSlots<T> F;
    int fp = 0, fe = 0;
    if constexpr (Pre != 0) {
        prow = xcd_block() * (kBlock / L) + (int)threadIdx.x / L;
        plane = (int)threadIdx.x % L;
        pvalid = prow < A.rows;
        if constexpr (Pre == 2) {
            if (pvalid) { fp = A.ptr[prow]; fe = A.ptr[prow + 1]; }
            fp += plane;
            flat_load<L>(A, fp, fe, F);
        } else {
            pre_slots<L>(A, prow, plane, pvalid, P);
        }
        if (pvalid && plane == 0) {
            py[0] = B.y[0][prow];
            py[1] = B.y[1][prow];
            if constexpr (!kDiff) {
                pax[0] = B.ax[0][prow];
                pax[1] = B.ax[1][prow];
            }
            pb = B.b[prow];
            ps = B.ysum[prow];
            pineq = B.ineq[prow];
        }
    }
    const DevState* st = S + p1;
    if (!st->active) return;
    const int cur = st->cur;
    const bool pending = st->pending != 0;
    const T eta_acc = (T)st->eta_acc;
    const T sigma = (T)(st->eta * st->omega);
    const T* __restrict__ xn = B.x[cur ^ 1];
    const T* __restrict__ y = B.y[cur];
    T* __restrict__ yn = B.y[cur ^ 1];
    const T* __restrict__ ax = B.ax[cur];
    T* __restrict__ axn = B.ax[cur ^ 1];
    double acc[2] = {0.0, 0.0};
    // FP32: A(xn - x) from the difference (see row_dot_diff); FP64: the
    // reference's 2 A xn - A x (pdlp_gpu_solve.cu:89) with A x carried.
    auto dot = [&](int row, int lane, bool valid, T (&v)[2]) {
        if constexpr (kDiff) {
            row_dot_diff<L>(A, reinterpret_cast<const float2*>(B.xd), row, lane, valid, v[0], v[1]);
        } else {
            v[0] = row_dot_pipe<L, T>(A, xn, row, lane, valid);
            v[1] = T(0);
        }
    };
    auto epilogue = [&](int row, T a, T ad) {
        axn[row] = a;
        const T yi = y[row];
        if (pending) B.ysum[row] += eta_acc * yi;
        // FP32: A(2 xn - x) = A xn + A (xn - x); FP64: 2 A xn - A x rounded
        // once, as the reference forms it (pdlp_gpu_solve.cu:89)
        T v;
        if constexpr (kDiff) {
            v = yi + sigma * (B.b[row] - (a + ad));
        } else {
            const T axr = ax[row];
            ad = a - axr;
            v = yi + sigma * (B.b[row] - (T(2) * a - axr));
        }
        if (B.ineq[row] && v < T(0)) v = T(0);
        yn[row] = v;
        const double d = (double)v - (double)yi;
        acc[0] += d * d;
        acc[1] += d * (double)ad;
    };
    if constexpr (Pre != 0) {
        T a, ad = T(0);
        if constexpr (Pre == 2) {
            if constexpr (kDiff)
                row_dot_diff_flat<L>(A, reinterpret_cast<const float2*>(B.xd), fp, fe, F, a, ad);
            else
                a = row_dot_flat<L, T>(A, xn, fp, fe, F);
        } else if constexpr (kDiff) {
            row_dot_diff_pre<L>(A, reinterpret_cast<const float2*>(B.xd), prow, P, a, ad);
        } else {
            a = row_dot_pre<L, true, T>(A, xn, prow, P);
        }
        if (pvalid && plane == 0) {
            // the epilogue above with the operands loaded before the state
            axn[prow] = a;
            const T yi = py[cur];
            if (pending) B.ysum[prow] = ps + eta_acc * yi;
            T v;
            if constexpr (kDiff) {
                v = yi + sigma * (pb - (a + ad));
            } else {
                const T axr = pax[cur];
                ad = a - axr;
                v = yi + sigma * (pb - (T(2) * a - axr));
            }
            if (pineq && v < T(0)) v = T(0);
            yn[prow] = v;
            const double d = (double)v - (double)yi;
            acc[0] += d * d;
            acc[1] += d * (double)ad;
        }
    } else if (staged_rows<L>(A.rows)) {
        for_row_tiles<L, 2, T>(A.rows, xcd_block(), gridDim.x, dot,
                               [&](int row, const T (&v)[2]) { epilogue(row, v[0], v[1]); });
    } else {
        for_rows<L>(A.rows, [&](int row, int lane, bool valid) {
            T v[2];
            dot(row, lane, valid, v);
            if (valid && lane == 0) epilogue(row, v[0], v[1]);
        });
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part_m[(size_t)blockIdx.x * 2] = acc[0];
        part_m[(size_t)blockIdx.x * 2 + 1] = acc[1];
    }
    if (blockIdx.x == gridDim.x - 1) {   // the last slot has the fewest rows
        double sn[1];
        sum_partials<1>(part_n, nb_n, sn);
        if (threadIdx.x == 0) *dx2_slot = sn[0];
    }
}

// KKT check, step 1 (blocks [0, en): columns; [en, en + em): rows): applies
// the last attempt's decision; if a check is due, adds the pending average
// term and forms the averages and the unscaled candidates
// xu = Dc x, xu2 = Dc x_avg, yu = Dr y, yu2 = Dr y_avg.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_check_prep(DevState* __restrict__ S, int n, int m,
                                                       int en, IterBufs<T> B, KktBufs K,
                                                       const double* __restrict__ part_m, int nb_m,
                                                       const double* __restrict__ dx2_slot) {
    __shared__ DevState Q;
    load_and_decide(S, 0, part_m, nb_m, dx2_slot, Q, blockIdx.x == 0);
    if (!Q.need_check) return;
    const int cur = Q.cur;
    const bool pending = Q.pending != 0;
    const T eta_acc = (T)Q.eta_acc;
    const double wsum = Q.wsum;
    if ((int)blockIdx.x < en) {
        for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += en * kBlock) {
            const T xj = B.x[cur][j];
            T xs = B.xsum[j];
            if (pending) { xs += eta_acc * xj; B.xsum[j] = xs; }
            const double avg = wsum > 0 ? (double)xs / wsum : (double)xj;
            K.xavg[j] = avg;
            K.xu[j] = K.Dc[j] * (double)xj;
            K.xu2[j] = K.Dc[j] * avg;
        }
    } else {
        const int b = blockIdx.x - en, em = gridDim.x - en;
        for (int i = b * kBlock + threadIdx.x; i < m; i += em * kBlock) {
            const T yi = B.y[cur][i];
            T ys = B.ysum[i];
            if (pending) { ys += eta_acc * yi; B.ysum[i] = ys; }
            const double avg = wsum > 0 ? (double)ys / wsum : (double)yi;
            K.yavg[i] = avg;
            K.yu[i] = K.Dr[i] * (double)yi;
            K.yu2[i] = K.Dr[i] * avg;
        }
    }
}

// KKT check, step 2: blocks [0, gkn) take the columns (rows of the unscaled
// A^T), blocks [gkn, ...) the rows of the unscaled A, both candidates at
// once (src/verify.cpp:24-94):
//   columns: bound violations, sign violations of z = c - A^T y, the column
//            part of the dual objective, c'x -> 8 partials per block;
//   rows:    row violations, sign violations of y, the row part of the
//            dual objective -> 6 partials per block.
// It also keeps the scaled product of the average A x_avg = Dr (A_u xu2),
// so a restart to the average needs no SpMV of A (A^T y is formed by every
// primal launch anyway).
template <int LA, int LAT>
__global__ __launch_bounds__(kBlock) void k_check_kkt(const DevState* __restrict__ S,
                                                      MatView<double> AuT, MatView<double> Au,
                                                      int gkn, KktBufs K,
                                                      double* __restrict__ part_kn,
                                                      double* __restrict__ part_km) {
    if (!S[1].need_check) return;
    if ((int)blockIdx.x < gkn) {
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const int G = gkn;
        const int gpb = kBlock / LAT;
        const int g = threadIdx.x / LAT, lane = threadIdx.x % LAT;
        for (int r0 = blockIdx.x * gpb; r0 < AuT.rows; r0 += G * gpb) {
            const int row = r0 + g;
            const bool valid = row < AuT.rows;
            double s1, s2;
            row_dot2<LAT>(AuT, K.yu, K.yu2, row, lane, valid, s1, s2);
            if (valid && lane == 0) {
                const double cj = K.c[row], lj = K.l[row], uj = K.u[row];
                const bool lf = isfinite(lj), uf = isfinite(uj);
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const double xv = k ? K.xu2[row] : K.xu[row];
                    const double z = cj - (k ? s2 : s1);
                    const double bv = smax(lj - xv, 0.0) + smax(xv - uj, 0.0);
                    double sv = 0.0;
                    if (!lf && z > 0) sv += z;
                    if (!uf && z < 0) sv += -z;
                    double zv = z, dcol = 0.0;
                    if (!lf && zv > 0) zv = 0.0;
                    if (!uf && zv < 0) zv = 0.0;
                    if (zv > 0 && lf) dcol += lj * zv;
                    if (zv < 0 && uf) dcol += -uj * -zv;
                    acc[4 * k + 0] += bv * bv;
                    acc[4 * k + 1] += sv * sv;
                    acc[4 * k + 2] += dcol;
                    acc[4 * k + 3] += cj * xv;
                }
            }
        }
        block_sum<8>(acc);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < 8; ++q) part_kn[(size_t)blockIdx.x * 8 + q] = acc[q];
        }
    } else {
        double acc[6] = {0, 0, 0, 0, 0, 0};
        const int G = gridDim.x - gkn, bb = blockIdx.x - gkn;
        const int gpb = kBlock / LA;
        const int g = threadIdx.x / LA, lane = threadIdx.x % LA;
        for (int r0 = bb * gpb; r0 < Au.rows; r0 += G * gpb) {
            const int row = r0 + g;
            const bool valid = row < Au.rows;
            double a1, a2;
            row_dot2<LA>(Au, K.xu, K.xu2, row, lane, valid, a1, a2);
            if (valid && lane == 0) {
                const double lo = K.rL[row], hi = K.rU[row];
                const bool lf = isfinite(lo), hf = isfinite(hi);
                K.ax_avg[row] = K.Dr[row] * a2;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const double a = k ? a2 : a1;
                    const double yv0 = k ? K.yu2[row] : K.yu[row];
                    const double rv = smax(lo - a, 0.0) + smax(a - hi, 0.0);
                    double sv = 0.0;
                    if (!lf && yv0 > 0) sv += yv0;
                    if (!hf && yv0 < 0) sv += -yv0;
                    double yv = yv0, drow = 0.0;
                    if (!lf && yv > 0) yv = 0.0;
                    if (!hf && yv < 0) yv = 0.0;
                    if (yv > 0 && lf) drow += lo * yv;
                    if (yv < 0 && hf) drow += -hi * -yv;
                    acc[3 * k + 0] += rv * rv;
                    acc[3 * k + 1] += sv * sv;
                    acc[3 * k + 2] += drow;
                }
            }
        }
        block_sum<6>(acc);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < 6; ++q) part_km[(size_t)bb * 6 + q] = acc[q];
        }
    }
}

// KKT check, step 3 (one block): eps of both candidates, keep-best,
// convergence, restart decision (include/pdlp_algo.hpp:135-174), and in the
// mixed mode the request to continue in FP64.
__global__ __launch_bounds__(kBlock) void k_kkt_fin(DevState* __restrict__ S,
                                                    const double* __restrict__ part_n, int nb_n,
                                                    const double* __restrict__ part_m, int nb_m) {
    DevState* st = S + 1;
    if (!st->need_check) return;
    double sn[8], sm[6];
    sum_partials<8>(part_n, nb_n, sn);
    sum_partials<6>(part_m, nb_m, sm);
    if (threadIdx.x == 0) kkt_fin_update(st, sn, sm);
}

// KKT check, step 4 (blocks [0, en): columns; the rest: rows): copy a new
// best point; on a restart measure |candidate - last restart point|
// (scaled), move the restart point, and when the candidate is the average
// move the iterate there together with its product A x_avg; clear the sums.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_check_restart(const DevState* __restrict__ S, int n,
                                                          int m, int en, IterBufs<T> B,
                                                          KktBufs K, double* __restrict__ part_rn,
                                                          double* __restrict__ part_rm) {
    const DevState* st = S + 1;
    if (!st->need_check) return;
    const int bw = st->best_which, rs = st->restart_now, ua = st->use_avg, cur = st->cur;
    double acc[1] = {0.0};
    if ((int)blockIdx.x < en) {
        for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += en * kBlock) {
            if (bw) K.best_x[j] = bw == 1 ? K.xu[j] : K.xu2[j];
            if (rs) {
                const double cand = ua ? K.xavg[j] : (double)B.x[cur][j];
                const double d = cand - (double)B.xlast[j];
                acc[0] += d * d;
                B.xlast[j] = (T)cand;
                if (ua) B.x[cur][j] = (T)cand;
                B.xsum[j] = T(0);
            }
        }
        block_sum<1>(acc);
        if (threadIdx.x == 0) part_rn[blockIdx.x] = acc[0];
    } else {
        const int b = blockIdx.x - en, em = gridDim.x - en;
        for (int i = b * kBlock + threadIdx.x; i < m; i += em * kBlock) {
            if (bw) K.best_y[i] = bw == 1 ? K.yu[i] : K.yu2[i];
            if (rs) {
                const double cand = ua ? K.yavg[i] : (double)B.y[cur][i];
                const double d = cand - (double)B.ylast[i];
                acc[0] += d * d;
                B.ylast[i] = (T)cand;
                if (ua) { B.y[cur][i] = (T)cand; B.ax[cur][i] = (T)K.ax_avg[i]; }
                B.ysum[i] = T(0);
            }
        }
        block_sum<1>(acc);
        if (threadIdx.x == 0) part_rm[b] = acc[0];
    }
}

// KKT check, step 5 (one block): primal-weight update and restart counters;
// resumes the iterations unless finished or switching precision; copies the
// state back to S[0] for the next replay (also when no check ran) and, when
// given, to the host-coherent mirror the host polls.
__global__ __launch_bounds__(kBlock) void k_restart_fin(DevState* __restrict__ S,
                                                        const double* __restrict__ part_n, int nb_n,
                                                        const double* __restrict__ part_m, int nb_m,
                                                        DevState* __restrict__ mirror) {
    DevState* st = S + 1;
    if (st->need_check) {
        double sn[1], sm[1];
        sum_partials<1>(part_n, nb_n, sn);
        sum_partials<1>(part_m, nb_m, sm);
        if (threadIdx.x == 0) restart_fin_update(st, sn[0], sm[0]);
    }
    if (threadIdx.x == 0) {
        st->groups += 1;
        S[0] = *st;
        if (mirror) *mirror = *st;   // host-coherent copy for the host's poll
    }
}

// Precision switch (mixed mode): the FP64 iterate starts at the FP32 loop's
// chosen candidate, as a restart (sums cleared, restart point moved).
__global__ __launch_bounds__(kBlock) void k_switch_n(const DevState* __restrict__ st, int n,
                                                     IterBufs<float> F, IterBufs<double> D,
                                                     KktBufs K) {
    const int ua = st->use_avg, cur = st->cur;
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const double cand = ua ? K.xavg[j] : (double)F.x[cur][j];
        D.x[0][j] = cand;
        D.xlast[j] = cand;
        D.xsum[j] = 0.0;
    }
}

__global__ __launch_bounds__(kBlock) void k_switch_m(const DevState* __restrict__ st, int m,
                                                     IterBufs<float> F, IterBufs<double> D,
                                                     KktBufs K) {
    const int ua = st->use_avg, cur = st->cur;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
        const double cand = ua ? K.yavg[i] : (double)F.y[cur][i];
        D.y[0][i] = cand;
        D.ylast[i] = cand;
        D.ysum[i] = 0.0;
    }
}

// ------------------------------------------------------- matrix layout --

// ELL slot k * rows + r <- CSR position r * w + k (every row has length w).
template <typename S, typename D>
__global__ __launch_bounds__(kBlock) void k_ell_from_csr(int32_t rows, int32_t w,
                                                         const S* __restrict__ src,
                                                         D* __restrict__ dst) {
    const int64_t total = (int64_t)rows * w;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * kBlock) {
        const int64_t r = q / w, k = q - r * w;
        dst[k * rows + r] = (D)src[q];
    }
}

// Sets *bad when some ELL slot q (row r) of M holds a value other than
// csc[idx[q]] * rsc[r], bit for bit -- the product the FP64 row products form
// instead of reading the value (MatView::rsc / csc).
__global__ __launch_bounds__(kBlock) void k_ell_rebuild_check(MatView<double> M,
                                                               const double* __restrict__ csc,
                                                               const double* __restrict__ rsc,
                                                               uint32_t* __restrict__ bad) {
    const int64_t total = (int64_t)M.rows * M.ell_w;
    uint32_t mine = 0;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < total;
         q += (int64_t)gridDim.x * kBlock) {
        const int64_t r = q % M.rows;   // slot-major: q = k * rows + r
        const double v = csc[M.idx[q]] * rsc[r];
        mine |= (uint32_t)(__double_as_longlong(v) != __double_as_longlong(M.val[q]));
    }
    if (mine) atomicOr(bad, 1u);
}

// Cold start of the device-resident loop in one launch: x = xlast = x0
// (rounded to T), y = ylast = 0, the averages' sums cleared, and, when S is
// given, both halves of the state set to st (passed by value).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_reset(int n, int m, const double* __restrict__ x0,
                                                  IterBufs<T> B, DevState* __restrict__ S,
                                                  DevState st) {
    const int stride = gridDim.x * kBlock;
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += stride) {
        const T v = (T)x0[j];
        B.x[0][j] = v;
        B.xlast[j] = v;
        B.xsum[j] = T(0);
    }
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += stride) {
        B.y[0][i] = T(0);
        B.ylast[i] = T(0);
        B.ysum[i] = T(0);
    }
    if (S && blockIdx.x == 0 && threadIdx.x == 0) {
        S[0] = st;
        S[1] = st;
    }
}

// out[0, len) = v (setup: vectors whose every entry is one value)
__global__ __launch_bounds__(kBlock) void k_fill(int64_t len, double v, double* __restrict__ out) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < len; q += (int64_t)gridDim.x * kBlock)
        out[q] = v;
}

// dst = (D) src, elementwise (FP64 -> FP32 rounds to nearest even).
template <typename S, typename D>
__global__ __launch_bounds__(kBlock) void k_convert(int64_t len, const S* __restrict__ src,
                                                    D* __restrict__ dst) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < len;
         q += (int64_t)gridDim.x * kBlock)
        dst[q] = (D)src[q];
}

}  // namespace supr
