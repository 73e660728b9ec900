// batch_kernels.hpp -- many independent LPs per launch: the device-resident
// PDHG loop of pdlp_kernels.hpp (include/pdlp_algo.hpp:49-185), segmented by
// instance.
//
// The LPs are concatenated block-diagonally: one CSR of A and one of A^T
// with instance-local row pointers and column indices, and concatenated
// vectors; BatchInst holds each instance's offsets.  Every launch covers all
// instances of the solve: instance k owns bn[k] consecutive blocks of each
// "column-side" launch and bm[k] of each "row-side" launch, found through
// the block -> instance tables.  Each instance has its own DevState pair,
// its own partial-sum slots and its own decisions (step size, checks,
// restarts, the FP32 -> FP64 switch), all taken with the single-LP code
// (decide_step, kkt_fin_update, restart_fin_update, switch_to_fp64).
//
// An instance's arithmetic depends only on its own data and on (bn, bm,
// lanes), which are functions of its shape and the engine configuration, so
// its result is the same bits alone and in any mix.  Instances that are
// paused for a check or finished return at block entry.
#pragma once
#include <type_traits>

#include "pdlp_device.hpp"

namespace supr {

struct BatchInst {
    int32_t m, n;        // PDLP form
    int32_t LA, LAT;     // lanes per row of A and of A^T (1, 4, 16, 64)
    int32_t bn, bm;      // blocks of this instance in column-side / row-side launches
    int32_t on, om;      // its first block in each (also its first partial slot)
    int32_t sn, sm;      // column / row side uses the LDS-staged tiles (for_row_tiles)
    int32_t nonneg;      // every l_j = +0 and u_j = +inf: the primal projects onto x >= 0
    int32_t id;          // the instance's slot (its state pair, its results)
    int64_t ro, co;      // offsets into the concatenated row / column vectors
    int64_t pa, pat;     // offsets of its row-pointer arrays (m + 1 and n + 1 entries)
    int64_t za;          // offset into the nonzero arrays (A and A^T hold nnz each)
};

// Shared description of one solve's launches.
struct BatchGeo {
    const BatchInst* inst;
    // block -> a copy of its instance's BatchInst (column side, row side): a
    // block reads its instance's description in one load instead of an
    // index and then the description
    const BatchInst* blk_n;
    const BatchInst* blk_m;
    const int32_t* ids;      // instances of this solve (one block each in the *_fin kernels)
    int32_t BN, BM;          // column-side / row-side blocks
    const int32_t *ptrA, *idxA, *ptrAT, *idxAT;
};

// Concatenated matrix values: scaled FP32 (mixed mode only), scaled FP64,
// unscaled FP64 (the KKT check).
struct BatchVals {
    const float *fA, *fAT;
    const double *dA, *dAT, *uA, *uAT;
};

template <typename V>
__device__ __forceinline__ MatView<V> inst_A(const BatchGeo& g, const BatchInst& I, const V* val) {
    return MatView<V>{I.m, 0, g.ptrA + I.pa, g.idxA + I.za, val + I.za};
}
template <typename V>
__device__ __forceinline__ MatView<V> inst_AT(const BatchGeo& g, const BatchInst& I, const V* val) {
    return MatView<V>{I.n, 0, g.ptrAT + I.pat, g.idxAT + I.za, val + I.za};
}

template <typename T>
__device__ __forceinline__ IterBufs<T> inst_bufs(const IterBufs<T>& B, const BatchInst& I) {
    IterBufs<T> r;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        r.x[q] = B.x[q] + I.co;
        r.y[q] = B.y[q] + I.ro;
        r.ax[q] = B.ax[q] + I.ro;
    }
    r.aty = B.aty + I.co;
    r.xsum = B.xsum + I.co;
    r.ysum = B.ysum + I.ro;
    r.xlast = B.xlast + I.co;
    r.ylast = B.ylast + I.ro;
    r.xd = B.xd ? B.xd + 2 * I.co : nullptr;
    r.c = B.c + I.co;
    r.l = B.l + I.co;
    r.u = B.u + I.co;
    r.b = B.b + I.ro;
    r.ineq = B.ineq + I.ro;
    return r;
}

__device__ __forceinline__ KktBufs inst_kkt(const KktBufs& K, const BatchInst& I) {
    KktBufs r;
    r.c = K.c + I.co; r.l = K.l + I.co; r.u = K.u + I.co;
    r.rL = K.rL + I.ro; r.rU = K.rU + I.ro;
    r.Dr = K.Dr + I.ro; r.Dc = K.Dc + I.co;
    r.xu = K.xu + I.co; r.xu2 = K.xu2 + I.co; r.xavg = K.xavg + I.co;
    r.yu = K.yu + I.ro; r.yu2 = K.yu2 + I.ro; r.yavg = K.yavg + I.ro;
    r.best_x = K.best_x + I.co; r.best_y = K.best_y + I.ro;
    r.aty_avg = K.aty_avg + I.co; r.ax_avg = K.ax_avg + I.ro;
    return r;
}

// a[i] for i in {0, 1} as a select (a runtime index into a local array
// would put the array in scratch memory).
template <typename P> __device__ __forceinline__ P sel(P const (&a)[2], int i) { return i ? a[1] : a[0]; }

// Which instance a block of a two-sided launch (column side first) serves.
struct Seat {
    BatchInst I; // the instance
    int k;       // its slot
    int lb, nb;  // block within the instance's side, blocks on that side
    bool col;    // column side
};
__device__ __forceinline__ Seat seat_two_sided(const BatchGeo& g) {
    Seat s;
    const int b = blockIdx.x;
    s.col = b < g.BN;
    s.I = s.col ? g.blk_n[b] : g.blk_m[b - g.BN];
    s.k = s.I.id;
    s.lb = s.col ? b - s.I.on : b - g.BN - s.I.om;
    s.nb = s.col ? s.I.bn : s.I.bm;
    return s;
}

// Rows of one instance, L lanes per row, its nb blocks striding; trip counts
// are block-uniform so group reductions see full groups.
template <int L, typename F>
__device__ __forceinline__ void inst_rows(int rows, int lb, int nb, F&& body) {
    constexpr int gpb = kBlock / L;
    const int g = threadIdx.x / L, lane = threadIdx.x % L;
    for (int r0 = lb * gpb; r0 < rows; r0 += nb * gpb) {
        const int row = r0 + g;
        body(row, lane, row < rows);
    }
}

// Runtime lanes-per-row dispatch (block-uniform); f(Lc) reads
// decltype(Lc)::value.
template <typename F> __device__ __forceinline__ void with_lanes(int L, F&& f) {
    switch (L) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 16: f(std::integral_constant<int, 16>{}); break;
        default: f(std::integral_constant<int, 64>{}); break;
    }
}

template <typename T> __device__ __forceinline__ const T* vals_A(const BatchVals& v);
template <> __device__ __forceinline__ const float* vals_A<float>(const BatchVals& v) { return v.fA; }
template <> __device__ __forceinline__ const double* vals_A<double>(const BatchVals& v) { return v.dA; }
template <typename T> __device__ __forceinline__ const T* vals_AT(const BatchVals& v);
template <> __device__ __forceinline__ const float* vals_AT<float>(const BatchVals& v) { return v.fAT; }
template <> __device__ __forceinline__ const double* vals_AT<double>(const BatchVals& v) { return v.dAT; }

// ---------------------------------------------------------- one attempt --

// Primal half (k_primal of pdlp_kernels.hpp) over one instance's columns.
template <typename T>
__device__ __forceinline__ double primal_rows(const BatchGeo& g, const BatchInst& I, int lb,
                                              const DevState& Q, const BatchVals& V,
                                              const IterBufs<T>& Bc) {
    const IterBufs<T> B = inst_bufs(Bc, I);
    const MatView<T> AT = inst_AT(g, I, vals_AT<T>(V));
    const int cur = Q.cur;
    const bool need_aty = Q.need_aty != 0;
    const bool pending = Q.pending != 0;
    const T eta_acc = (T)Q.eta_acc;
    const T tau = (T)(Q.eta / Q.omega);
    const T* __restrict__ x = sel(B.x, cur);
    T* __restrict__ xn = sel(B.x, cur ^ 1);
    const T* __restrict__ y = sel(B.y, cur);
    double acc = 0.0;
    const bool nonneg = I.nonneg != 0;
    auto epilogue = [&](int row, T a) {
        const T xj = x[row];
        if (pending) B.xsum[row] += eta_acc * xj;
        const T z = xj - tau * (B.c[row] - a);
        // smin(smax(z, +0), +inf) is smax(z, +0) bit for bit: no l / u reads
        const T v = nonneg ? smax(z, T(0)) : smin(smax(z, B.l[row]), B.u[row]);
        xn[row] = v;
        if constexpr (sizeof(T) == 4)
            reinterpret_cast<float2*>(B.xd)[row] = make_float2(v, v - xj);
        const double d = (double)v - (double)xj;
        acc += d * d;
    };
    with_lanes(I.LAT, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        if (I.sn) {
            if (need_aty) {
                for_row_tiles<L, 1, T>(
                    I.n, lb, I.bn,
                    [&](int row, int lane, bool valid, T (&v)[1]) { v[0] = row_dot<L, T>(AT, y, row, lane, valid); },
                    [&](int row, const T (&v)[1]) {
                        B.aty[row] = v[0];
                        epilogue(row, v[0]);
                    });
            } else {
                for_rows_flat(I.n, lb, I.bn, [&](int row) { epilogue(row, B.aty[row]); });
            }
            return;
        }
        inst_rows<L>(I.n, lb, I.bn, [&](int row, int lane, bool valid) {
            T a = T(0);
            if (need_aty) {
                a = row_dot<L, T>(AT, y, row, lane, valid);
                if (valid && lane == 0) B.aty[row] = a;
            }
            if (valid && lane == 0) {
                if (!need_aty) a = B.aty[row];
                epilogue(row, a);
            }
        });
    });
    return acc;
}

// Column side of an attempt.  S: DevState pairs (S[2k], S[2k+1]); p: the
// graph slot parity.  Every block of instance k applies k's pending step
// decision from its dual partials; k's first block publishes it.
__global__ __launch_bounds__(kBlock) void kb_primal(DevState* __restrict__ S, int p, BatchGeo g,
                                                    BatchVals V, IterBufs<float> F,
                                                    IterBufs<double> D,
                                                    const double* __restrict__ part_m,
                                                    const double* __restrict__ dx2,
                                                    double* __restrict__ part_n) {
    __shared__ DevState Q;
    const int b = blockIdx.x;
    const BatchInst I = g.blk_n[b];
    const int k = I.id;
    const int lb = b - I.on;
    DevState* Sk = S + 2 * k;
    // the publisher is the instance's last block (the fewest rows): the step
    // powers it forms before its rows stay off the launch's critical path
    const bool publisher = lb == I.bn - 1;
    load_and_decide(Sk, p, part_m + 2 * (size_t)I.om, I.bm, dx2 + k, Q, publisher);
    if (!Q.active) return;
    if (publisher && threadIdx.x == 0) {
        Sk[p ^ 1].decide_pending = 1;   // kb_dual will run
        publish_step_powers(Sk, p, Q);
    }
    double acc[1];
    acc[0] = Q.phase == 0 ? primal_rows<float>(g, I, lb, Q, V, F) : primal_rows<double>(g, I, lb, Q, V, D);
    block_sum<1>(acc);
    if (threadIdx.x == 0) part_n[b] = acc[0];
}

// Dual half (k_dual of pdlp_kernels.hpp) over one instance's rows.
template <typename T>
__device__ __forceinline__ void dual_rows(const BatchGeo& g, const BatchInst& I, int lb,
                                          const DevState* st, const BatchVals& V,
                                          const IterBufs<T>& Bc, double (&acc)[2]) {
    const IterBufs<T> B = inst_bufs(Bc, I);
    const MatView<T> A = inst_A(g, I, vals_A<T>(V));
    const int cur = st->cur;
    const bool pending = st->pending != 0;
    const T eta_acc = (T)st->eta_acc;
    const T sigma = (T)(st->eta * st->omega);
    const T* __restrict__ xn = sel(B.x, cur ^ 1);
    const T* __restrict__ y = sel(B.y, cur);
    T* __restrict__ yn = sel(B.y, cur ^ 1);
    const T* __restrict__ ax = sel(B.ax, cur);
    T* __restrict__ axn = sel(B.ax, cur ^ 1);
    constexpr bool kDiff = sizeof(T) == 4;
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
    with_lanes(I.LA, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        auto dot = [&](int row, int lane, bool valid, T (&v)[2]) {
            if constexpr (kDiff) {
                row_dot_diff<L>(A, reinterpret_cast<const float2*>(B.xd), row, lane, valid, v[0], v[1]);
            } else {
                v[0] = row_dot<L, T>(A, xn, row, lane, valid);
                v[1] = T(0);
            }
        };
        if (I.sm) {
            for_row_tiles<L, 2, T>(I.m, lb, I.bm, dot,
                                   [&](int row, const T (&v)[2]) { epilogue(row, v[0], v[1]); });
            return;
        }
        inst_rows<L>(I.m, lb, I.bm, [&](int row, int lane, bool valid) {
            T v[2];
            dot(row, lane, valid, v);
            if (valid && lane == 0) epilogue(row, v[0], v[1]);
        });
    });
}

// Row side of an attempt; the last block of each instance also sums its
// primal partials into dx2[k].
__global__ __launch_bounds__(kBlock) void kb_dual(const DevState* __restrict__ S, int p1,
                                                  BatchGeo g, BatchVals V, IterBufs<float> F,
                                                  IterBufs<double> D,
                                                  const double* __restrict__ part_n,
                                                  double* __restrict__ dx2,
                                                  double* __restrict__ part_m) {
    const int b = blockIdx.x;
    const BatchInst I = g.blk_m[b];
    const int k = I.id;
    const int lb = b - I.om;
    const DevState* st = S + 2 * k + p1;
    if (!st->active) return;
    double acc[2] = {0.0, 0.0};
    if (st->phase == 0) dual_rows<float>(g, I, lb, st, V, F, acc);
    else dual_rows<double>(g, I, lb, st, V, D, acc);
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        part_m[(size_t)b * 2] = acc[0];
        part_m[(size_t)b * 2 + 1] = acc[1];
    }
    if (lb == I.bm - 1) {   // the instance's last block has the fewest rows
        double sn[1];
        sum_partials<1>(part_n + I.on, I.bn, sn);
        if (threadIdx.x == 0) dx2[k] = sn[0];
    }
}

// ------------------------------------------------------- the KKT check --

template <typename T>
__device__ __forceinline__ void prep_elems(const Seat& s, const BatchInst& I, const DevState& Q,
                                           const IterBufs<T>& Bc, const KktBufs& Kc) {
    const IterBufs<T> B = inst_bufs(Bc, I);
    const KktBufs K = inst_kkt(Kc, I);
    const int cur = Q.cur;
    const bool pending = Q.pending != 0;
    const T eta_acc = (T)Q.eta_acc;
    const double wsum = Q.wsum;
    if (s.col) {
        for (int j = s.lb * kBlock + threadIdx.x; j < I.n; j += s.nb * kBlock) {
            const T xj = sel(B.x, cur)[j];
            T xs = B.xsum[j];
            if (pending) { xs += eta_acc * xj; B.xsum[j] = xs; }
            const double avg = wsum > 0 ? (double)xs / wsum : (double)xj;
            K.xavg[j] = avg;
            K.xu[j] = K.Dc[j] * (double)xj;
            K.xu2[j] = K.Dc[j] * avg;
        }
    } else {
        for (int i = s.lb * kBlock + threadIdx.x; i < I.m; i += s.nb * kBlock) {
            const T yi = sel(B.y, cur)[i];
            T ys = B.ysum[i];
            if (pending) { ys += eta_acc * yi; B.ysum[i] = ys; }
            const double avg = wsum > 0 ? (double)ys / wsum : (double)yi;
            K.yavg[i] = avg;
            K.yu[i] = K.Dr[i] * (double)yi;
            K.yu2[i] = K.Dr[i] * avg;
        }
    }
}

// Check step 1 (k_check_prep): apply the last decision (S[2k] -> S[2k+1]);
// if a check is due, averages and unscaled candidates.
__global__ __launch_bounds__(kBlock) void kb_check_prep(DevState* __restrict__ S, BatchGeo g,
                                                        IterBufs<float> F, IterBufs<double> D,
                                                        KktBufs K,
                                                        const double* __restrict__ part_m,
                                                        const double* __restrict__ dx2) {
    __shared__ DevState Q;
    const Seat s = seat_two_sided(g);
    const BatchInst& I = s.I;
    load_and_decide(S + 2 * s.k, 0, part_m + 2 * (size_t)I.om, I.bm, dx2 + s.k, Q,
                    s.col && s.lb == 0);
    if (!Q.need_check) return;
    if (Q.phase == 0) prep_elems<float>(s, I, Q, F, K);
    else prep_elems<double>(s, I, Q, D, K);
}

// Check step 2 (k_check_kkt): column side -> 8 partials per block, row side
// -> 6, both candidates at once (src/verify.cpp:24-94); keeps A^T y_avg and
// A x_avg for a restart to the average.
__global__ __launch_bounds__(kBlock) void kb_check_kkt(const DevState* __restrict__ S, BatchGeo g,
                                                       BatchVals V, KktBufs Kc,
                                                       double* __restrict__ part_kn,
                                                       double* __restrict__ part_km) {
    const Seat s = seat_two_sided(g);
    const BatchInst& I = s.I;
    if (!S[2 * s.k + 1].need_check) return;
    const KktBufs K = inst_kkt(Kc, I);
    if (s.col) {
        const MatView<double> AuT = inst_AT(g, I, V.uAT);
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        with_lanes(I.LAT, [&](auto Lc) {
            constexpr int L = decltype(Lc)::value;
            inst_rows<L>(I.n, s.lb, s.nb, [&](int row, int lane, bool valid) {
                double s1, s2;
                row_dot2<L>(AuT, K.yu, K.yu2, row, lane, valid, s1, s2);
                if (valid && lane == 0) {
                    const double cj = K.c[row], lj = K.l[row], uj = K.u[row];
                    const bool lf = isfinite(lj), uf = isfinite(uj);
                    K.aty_avg[row] = K.Dc[row] * s2;
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const double xv = q ? K.xu2[row] : K.xu[row];
                        const double z = cj - (q ? s2 : s1);
                        const double bv = smax(lj - xv, 0.0) + smax(xv - uj, 0.0);
                        double sv = 0.0;
                        if (!lf && z > 0) sv += z;
                        if (!uf && z < 0) sv += -z;
                        double zv = z, dcol = 0.0;
                        if (!lf && zv > 0) zv = 0.0;
                        if (!uf && zv < 0) zv = 0.0;
                        if (zv > 0 && lf) dcol += lj * zv;
                        if (zv < 0 && uf) dcol += -uj * -zv;
                        acc[4 * q + 0] += bv * bv;
                        acc[4 * q + 1] += sv * sv;
                        acc[4 * q + 2] += dcol;
                        acc[4 * q + 3] += cj * xv;
                    }
                }
            });
        });
        block_sum<8>(acc);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < 8; ++q) part_kn[(size_t)blockIdx.x * 8 + q] = acc[q];
        }
    } else {
        const MatView<double> Au = inst_A(g, I, V.uA);
        double acc[6] = {0, 0, 0, 0, 0, 0};
        with_lanes(I.LA, [&](auto Lc) {
            constexpr int L = decltype(Lc)::value;
            inst_rows<L>(I.m, s.lb, s.nb, [&](int row, int lane, bool valid) {
                double a1, a2;
                row_dot2<L>(Au, K.xu, K.xu2, row, lane, valid, a1, a2);
                if (valid && lane == 0) {
                    const double lo = K.rL[row], hi = K.rU[row];
                    const bool lf = isfinite(lo), hf = isfinite(hi);
                    K.ax_avg[row] = K.Dr[row] * a2;
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const double a = q ? a2 : a1;
                        const double yv0 = q ? K.yu2[row] : K.yu[row];
                        const double rv = smax(lo - a, 0.0) + smax(a - hi, 0.0);
                        double sv = 0.0;
                        if (!lf && yv0 > 0) sv += yv0;
                        if (!hf && yv0 < 0) sv += -yv0;
                        double yv = yv0, drow = 0.0;
                        if (!lf && yv > 0) yv = 0.0;
                        if (!hf && yv < 0) yv = 0.0;
                        if (yv > 0 && lf) drow += lo * yv;
                        if (yv < 0 && hf) drow += -hi * -yv;
                        acc[3 * q + 0] += rv * rv;
                        acc[3 * q + 1] += sv * sv;
                        acc[3 * q + 2] += drow;
                    }
                }
            });
        });
        block_sum<6>(acc);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < 6; ++q) part_km[(size_t)(blockIdx.x - g.BN) * 6 + q] = acc[q];
        }
    }
}

// Check step 3 (k_kkt_fin), one block per instance of the solve.
__global__ __launch_bounds__(kBlock) void kb_kkt_fin(DevState* __restrict__ S, BatchGeo g,
                                                     const double* __restrict__ part_kn,
                                                     const double* __restrict__ part_km) {
    const int k = g.ids[blockIdx.x];
    const BatchInst I = g.inst[k];
    DevState* st = S + 2 * k + 1;
    if (!st->need_check) return;
    double sn[8], sm[6];
    sum_partials<8>(part_kn + 8 * (size_t)I.on, I.bn, sn);
    sum_partials<6>(part_km + 6 * (size_t)I.om, I.bm, sm);
    if (threadIdx.x == 0) kkt_fin_update(st, sn, sm);
}

// Check step 4 (k_check_restart): best copy; restart distance, restart
// point, move to the average with its products; clear the sums.  In the
// mixed mode an instance that asked to continue in FP64 gets its FP64
// iterate here (the candidate, as a restart: k_switch_n / k_switch_m).
template <typename T>
__device__ __forceinline__ double restart_elems(const Seat& s, const BatchInst& I,
                                                const DevState* st, const IterBufs<T>& Bc,
                                                const IterBufs<double>& Dc, const KktBufs& Kc) {
    const IterBufs<T> B = inst_bufs(Bc, I);
    const KktBufs K = inst_kkt(Kc, I);
    const int bw = st->best_which, rs = st->restart_now, ua = st->use_avg, cur = st->cur;
    const bool sw = sizeof(T) == 4 && st->switch_req != 0;
    double acc = 0.0;
    if (s.col) {
        for (int j = s.lb * kBlock + threadIdx.x; j < I.n; j += s.nb * kBlock) {
            if (bw) K.best_x[j] = bw == 1 ? K.xu[j] : K.xu2[j];
            if (rs) {
                const double cand = ua ? K.xavg[j] : (double)sel(B.x, cur)[j];
                const double d = cand - (double)B.xlast[j];
                acc += d * d;
                B.xlast[j] = (T)cand;
                if (ua) { sel(B.x, cur)[j] = (T)cand; B.aty[j] = (T)K.aty_avg[j]; }
                B.xsum[j] = T(0);
            }
            if (sw) {
                const double cand = ua ? K.xavg[j] : (double)sel(B.x, cur)[j];
                Dc.x[0][I.co + j] = cand;
                Dc.xlast[I.co + j] = cand;
                Dc.xsum[I.co + j] = 0.0;
            }
        }
    } else {
        for (int i = s.lb * kBlock + threadIdx.x; i < I.m; i += s.nb * kBlock) {
            if (bw) K.best_y[i] = bw == 1 ? K.yu[i] : K.yu2[i];
            if (rs) {
                const double cand = ua ? K.yavg[i] : (double)sel(B.y, cur)[i];
                const double d = cand - (double)B.ylast[i];
                acc += d * d;
                B.ylast[i] = (T)cand;
                if (ua) { sel(B.y, cur)[i] = (T)cand; sel(B.ax, cur)[i] = (T)K.ax_avg[i]; }
                B.ysum[i] = T(0);
            }
            if (sw) {
                const double cand = ua ? K.yavg[i] : (double)sel(B.y, cur)[i];
                Dc.y[0][I.ro + i] = cand;
                Dc.ylast[I.ro + i] = cand;
                Dc.ysum[I.ro + i] = 0.0;
            }
        }
    }
    return acc;
}

__global__ __launch_bounds__(kBlock) void kb_check_restart(const DevState* __restrict__ S,
                                                           BatchGeo g, IterBufs<float> F,
                                                           IterBufs<double> D, KktBufs K,
                                                           double* __restrict__ part_rn,
                                                           double* __restrict__ part_rm) {
    const Seat s = seat_two_sided(g);
    const BatchInst& I = s.I;
    const DevState* st = S + 2 * s.k + 1;
    if (!st->need_check) return;
    double acc[1];
    acc[0] = st->phase == 0 ? restart_elems<float>(s, I, st, F, D, K)
                            : restart_elems<double>(s, I, st, D, D, K);
    block_sum<1>(acc);
    if (threadIdx.x == 0) {
        if (s.col) part_rn[blockIdx.x] = acc[0];
        else part_rm[blockIdx.x - g.BN] = acc[0];
    }
}

// Mixed mode, after kb_check_restart: A^T y and A x (FP64) of the new FP64
// iterate of each instance that switches (set_products of the single-LP
// engine).
__global__ __launch_bounds__(kBlock) void kb_switch_products(const DevState* __restrict__ S,
                                                             BatchGeo g, BatchVals V,
                                                             IterBufs<double> Dc) {
    const Seat s = seat_two_sided(g);
    const BatchInst& I = s.I;
    const DevState* st = S + 2 * s.k + 1;
    if (!st->need_check || !st->switch_req) return;
    const IterBufs<double> B = inst_bufs(Dc, I);
    if (s.col) {
        const MatView<double> AT = inst_AT(g, I, V.dAT);
        with_lanes(I.LAT, [&](auto Lc) {
            constexpr int L = decltype(Lc)::value;
            inst_rows<L>(I.n, s.lb, s.nb, [&](int row, int lane, bool valid) {
                const double a = row_dot<L, double>(AT, B.y[0], row, lane, valid);
                if (valid && lane == 0) B.aty[row] = a;
            });
        });
    } else {
        const MatView<double> A = inst_A(g, I, V.dA);
        with_lanes(I.LA, [&](auto Lc) {
            constexpr int L = decltype(Lc)::value;
            inst_rows<L>(I.m, s.lb, s.nb, [&](int row, int lane, bool valid) {
                const double a = row_dot<L, double>(A, B.x[0], row, lane, valid);
                if (valid && lane == 0) B.ax[0][row] = a;
            });
        });
    }
}

// Check step 5 (k_restart_fin), one block per instance: primal weight and
// restart counters, the FP64 hand-over, S[2k] = S[2k+1] for the next replay,
// and the instance's done flag for the host.
__global__ __launch_bounds__(kBlock) void kb_restart_fin(DevState* __restrict__ S, BatchGeo g,
                                                         const double* __restrict__ part_rn,
                                                         const double* __restrict__ part_rm,
                                                         int32_t* __restrict__ done_flags) {
    const int k = g.ids[blockIdx.x];
    const BatchInst I = g.inst[k];
    DevState* st = S + 2 * k + 1;
    if (st->need_check) {
        double sn[1], sm[1];
        sum_partials<1>(part_rn + I.on, I.bn, sn);
        sum_partials<1>(part_rm + I.om, I.bm, sm);
        if (threadIdx.x == 0) {
            restart_fin_update(st, sn[0], sm[0]);
            if (st->switch_req) switch_to_fp64(*st);
        }
    }
    if (threadIdx.x == 0) {
        S[2 * k] = *st;
        if (done_flags) done_flags[k] = st->done;   // host-coherent, polled by the server
    }
}

// After each replay: how many instances of the solve are still running, and
// a progress count (attempts + checks summed), for the host's poll.
__global__ __launch_bounds__(kBlock) void kb_status(const DevState* __restrict__ S, BatchGeo g,
                                                    int count, double* __restrict__ out) {
    double acc[2] = {0.0, 0.0};
    for (int r = threadIdx.x; r < count; r += kBlock) {
        const DevState& st = S[2 * g.ids[r]];
        acc[0] += st.done ? 0.0 : 1.0;
        acc[1] += (double)st.attempts + (double)st.checks;
    }
    block_sum<2>(acc);
    if (threadIdx.x == 0) {
        out[0] = acc[0];
        out[1] = acc[1];
    }
}

// Cold start of the instances of a solve in their starting precision:
// x = clip(0, l, u), y = 0, restart points = the start, sums = 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void kb_cold(BatchGeo g, IterBufs<T> Bc) {
    const Seat s = seat_two_sided(g);
    const BatchInst& I = s.I;
    const IterBufs<T> B = inst_bufs(Bc, I);
    if (s.col) {
        for (int j = s.lb * kBlock + threadIdx.x; j < I.n; j += s.nb * kBlock) {
            const T v = smin(smax(T(0), B.l[j]), B.u[j]);
            B.x[0][j] = v;
            B.xlast[j] = v;
            B.xsum[j] = T(0);
        }
    } else {
        for (int i = s.lb * kBlock + threadIdx.x; i < I.m; i += s.nb * kBlock) {
            B.y[0][i] = T(0);
            B.ylast[i] = T(0);
            B.ysum[i] = T(0);
        }
    }
}

// A x of the cold start (row side only; grid = BM blocks).
template <typename T>
__global__ __launch_bounds__(kBlock) void kb_cold_ax(BatchGeo g, BatchVals V, IterBufs<T> Bc) {
    const BatchInst I = g.blk_m[blockIdx.x];
    const IterBufs<T> B = inst_bufs(Bc, I);
    const MatView<T> A = inst_A(g, I, vals_A<T>(V));
    with_lanes(I.LA, [&](auto Lc) {
        constexpr int L = decltype(Lc)::value;
        inst_rows<L>(I.m, blockIdx.x - I.om, I.bm, [&](int row, int lane, bool valid) {
            const T a = row_dot<L, T>(A, B.x[0], row, lane, valid);
            if (valid && lane == 0) B.ax[0][row] = a;
        });
    });
}

// Initial states of newly started instances: S[2k] = S[2k+1] = src[r] for
// k = slots[r].
__global__ __launch_bounds__(kBlock) void kb_put_states(DevState* __restrict__ S,
                                                        const DevState* __restrict__ src,
                                                        const int32_t* __restrict__ slots,
                                                        int count) {
    for (int r = blockIdx.x * kBlock + threadIdx.x; r < count; r += gridDim.x * kBlock) {
        const DevState s = src[r];
        S[2 * slots[r]] = s;
        S[2 * slots[r] + 1] = s;
    }
}

}  // namespace supr
