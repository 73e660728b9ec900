// pdlp_device.hpp -- device helpers and layouts shared by the single-LP
// engine (pdlp_kernels.hpp) and the batched engine (batch_kernels.hpp):
// CDNA4 (gfx950, wave64) reductions, SpMV row products, control structs.
//
// Two families:
//  * plain kernels (k_spmv, k_plain_*): one operation per launch, the shape of
//    the reference GPU backend (gpu/pdlp_gpu_solve.cu:42-138) -- used by the
//    host-driven plain FP64 mode and by the SpMV autotuner;
//  * device-resident kernels (k_primal, k_dual, k_check_*, k_*_fin): the
//    whole PDHG loop of include/pdlp_algo.hpp:103-175 with its step-size rule,
//    ergodic averages, KKT check and restart rule evaluated on the device.
//    Control flow lives in a DevState block in device memory; each kernel
//    reads it at entry and returns at once when it has nothing to do, so a
//    fixed sequence of launches (a replayed hipGraph) can run a variable
//    number of accepted / rejected steps and checks without a host round trip.
//
// SpMV schedules.  A row is handled by a group of L lanes: L = 1, 2, ..., 64
// lanes of one wavefront (lanes stride the row's nonzeros, an xor-shuffle
// tree combines them), or L = 256, the whole block (the LDS-staged row split
// for long rows: four wavefront sums combined through LDS).  Matrices whose rows all have the same short
// length (A^T of a transportation LP: 2 per row) are stored as ELL instead
// (column-major slots, coalesced across the wavefront, no row pointers).  The
// engine picks L per matrix by timing the candidates (autotune) or from the
// mean row length.
//
// Reductions are deterministic: each block writes its partial sums and a
// one-block kernel adds them in a fixed order.  The hot kernels carry no
// atomics and no fences: kernel boundaries order everything.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

namespace supr {

constexpr int kBlock = 256;       // 4 wavefronts
constexpr int kWaves = kBlock / 64;
constexpr int kMaxBlocks = 2048;  // grid-stride cap: 8 blocks per CU on 256 CUs

// std::max / std::min semantics (ties and signed zeros as libstdc++).
template <typename T> __host__ __device__ __forceinline__ T smax(T a, T b) { return a < b ? b : a; }
template <typename T> __host__ __device__ __forceinline__ T smin(T a, T b) { return b < a ? b : a; }

// Control block of the device-resident loop (one per LP).
struct DevState {
    // step size, primal weight, averaging
    double eta, omega, wsum, eta_acc;
    // restart bookkeeping
    double merit_last_restart, merit_prev_cand, cand_merit;
    double best_merit, stall_ref;
    double best_eps[3];
    double last_eps[3];      // eps of the candidate chosen at the last check
    double last_obj;
    double best_obj;         // c'x + offset of the point kept as best (unscaled form)
    double tol, switch_tol;
    double denom_p, denom_d, obj_offset;
    // (accepted + 2)^-0.3 and ^-0.6 for the next step decision, formed by a
    // block of the previous primal launch off the decision's critical path;
    // valid when pw_k1 == accepted + 2 (otherwise decide_step calls pow)
    double pw3, pw6;
    int32_t attempts, accepted, since_restart, restarts, checks;
    int32_t max_iter, check_every, stall_checks, stall;
    int32_t cur;             // which half of each double buffer is current
    int32_t active;          // iterations run (0: paused for a check, or finished)
    int32_t need_aty;        // next k_primal recomputes A^T y
    int32_t decide_pending;  // the last attempt's step decision is not applied yet
    int32_t pending;         // last accepted step not yet added to the averages
    int32_t need_check;
    int32_t done;            // 1 converged, 2 iteration limit
    int32_t has_best, best_which, use_avg, restart_now;
    int32_t switch_req, phase;   // switch_req: reason code (1 merit, 2 stall near floor, 3 long stall); phase 0: FP32 iterates, 1: FP64
    int32_t attempts_fp32;   // attempts when the FP32 loop handed over (mixed mode)
    int32_t pw_k1;           // the k1 that pw3 / pw6 belong to (0: none)
    int32_t groups;          // check groups run (k_restart_fin), for the host's launch plan
};

template <typename V> struct MatView {
    int32_t rows;
    int32_t ell_w;           // > 0: ELL with this width (idx / val column-major)
    const int32_t* ptr;      // CSR row pointers (unused for ELL)
    const int32_t* idx;
    const V* val;
    // ELL only: the same indices as 16-bit values when the gathered vector
    // has at most 65,536 entries (A^T of an LP with few rows); null otherwise.
    // The ELL loops read these instead of idx: 2 bytes per slot instead of 4.
    const uint16_t* idx16 = nullptr;
    // ELL, FP64 values only: when every stored value of row r, slot k equals
    // csc[idx] * rsc[r] bit for bit (a matrix whose unscaled entries share one
    // value a: scaled a_ij = (Dr_i a) Dc_j), row_dot forms each value from
    // these two factors instead of reading it -- one factor per row and one
    // gathered from a short, cached vector instead of 8 bytes per slot.  The
    // products are the stored values, so the sums are the same bits.
    const double* rsc = nullptr;
    const double* csc = nullptr;
    // ELL with one lane per row, an even row count and more rows than one
    // grid-stride pass holds: k_primal gives each lane two adjacent rows
    // (ell_pair_dot), so its row streams move 8 or 16 bytes per lane access.
    int32_t pairs = 0;
};

// Iterate buffers of one precision.
template <typename T> struct IterBufs {
    T* x[2];
    T* y[2];
    T* ax[2];
    T* aty;
    T* xsum;
    T* ysum;
    T* xlast;
    T* ylast;
    T* xd;        // FP32 loop: interleaved pairs (xn, xn - x) for k_dual's gather
    const T* c;
    const T* l;   // l, u null: every l_j = +0 and u_j = +inf (x >= 0)
    const T* u;
    const T* b;
    const uint8_t* ineq;
};

// Vectors of the FP64 KKT check (unscaled PDLP form).
struct KktBufs {
    const double* c;   // unscaled c, l, u (n) and rL, rU (m)
    const double* l;
    const double* u;
    const double* rL;
    const double* rU;
    const double* Dr;
    const double* Dc;
    double* xu;        // Dc * x, Dc * x_avg
    double* xu2;
    double* xavg;      // scaled average
    double* yu;
    double* yu2;
    double* yavg;
    double* best_x;
    double* best_y;
    double* aty_avg;   // scaled A^T y_avg (n), A x_avg (m) from the check
    double* ax_avg;
};

// ------------------------------------------------------------- helpers --

// The value held by lane (lane ^ O) of the wavefront, O a power of two below
// 64, moved without LDS: DPP inside a row of 16 lanes (quad_perm for 1 and 2,
// row_shl / row_shr for 4, row_ror for 8) and the permlane swaps across rows
// (16) and wavefront halves (32).  __shfl_xor goes through ds_bpermute, an
// LDS round trip per 32-bit half.  All lanes of the wavefront must be active.
template <int O> __device__ __forceinline__ int xor_lane(int v) {
    if constexpr (O == 1) {
        return __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xf, 0xf, false);   // quad_perm [1,0,3,2]
    } else if constexpr (O == 2) {
        return __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xf, 0xf, false);   // quad_perm [2,3,0,1]
    } else if constexpr (O == 4) {
        const int up = __builtin_amdgcn_update_dpp(v, v, 0x104, 0xf, 0xf, false);   // row_shl:4
        const int dn = __builtin_amdgcn_update_dpp(v, v, 0x114, 0xf, 0xf, false);   // row_shr:4
        return (threadIdx.x & 4) ? dn : up;
    } else if constexpr (O == 8) {
        return __builtin_amdgcn_update_dpp(v, v, 0x128, 0xf, 0xf, false);  // row_ror:8
    } else if constexpr (O == 16) {
        // odd rows of the first operand <-> even rows of the second
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (threadIdx.x & 16) ? r[0] : r[1];
    } else {
        static_assert(O == 32, "xor distances 1..32");
        // upper half of the first operand <-> lower half of the second
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return (threadIdx.x & 32) ? r[0] : r[1];
    }
}
template <int O> __device__ __forceinline__ float xor_lane(float v) {
    return __int_as_float(xor_lane<O>(__float_as_int(v)));
}
template <int O> __device__ __forceinline__ double xor_lane(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)xor_lane<O>((int)(unsigned)b);
    const unsigned hi = (unsigned)xor_lane<O>((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Butterfly sum over groups of W lanes (W = 2..64): s += s(lane ^ o) for
// o = W/2, ..., 1 -- the additions of a __shfl_xor butterfly, in the same
// order, so the same bits; every lane of a group ends with the group's sum.
template <int W, typename T> __device__ __forceinline__ T xor_sum(T s) {
    if constexpr (W >= 64) s += xor_lane<32>(s);
    if constexpr (W >= 32) s += xor_lane<16>(s);
    if constexpr (W >= 16) s += xor_lane<8>(s);
    if constexpr (W >= 8) s += xor_lane<4>(s);
    if constexpr (W >= 4) s += xor_lane<2>(s);
    if constexpr (W >= 2) s += xor_lane<1>(s);
    return s;
}

// Wavefront sum, valid in every lane.  Lane 0's value is that of the
// __shfl_down tree (v_i + v_{i+32}, then + the i+16 partner, ...): for the
// lanes below each distance the xor partner is lane + o.
__device__ __forceinline__ double wave_sum(double v) { return xor_sum<64>(v); }

// Sums N values over the block; result valid in thread 0.
template <int N> __device__ __forceinline__ void block_sum(double (&v)[N]) {
    __shared__ double lds[N][kWaves];
#pragma unroll
    for (int q = 0; q < N; ++q) v[q] = wave_sum(v[q]);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int q = 0; q < N; ++q) lds[q][w] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < N; ++q) {
            double s = lds[q][0];
            for (int k = 1; k < kWaves; ++k) s += lds[q][k];
            v[q] = s;
        }
    }
    __syncthreads();
}

// Deterministic sum of N-wide partials of nb <= kMaxBlocks blocks by one
// block (a separate launch, so the producers' stores are visible): thread t
// loads blocks t, t+256, ... with independent loads, adds them in that order,
// then a fixed tree.  Valid in thread 0.
template <int N>
__device__ __forceinline__ void sum_partials(const double* __restrict__ part, int nb,
                                             double (&out)[N]) {
    constexpr int K = kMaxBlocks / kBlock;
    double v[K][N];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int b = threadIdx.x + k * kBlock;
#pragma unroll
        for (int q = 0; q < N; ++q) v[k][q] = b < nb ? part[(size_t)b * N + q] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < N; ++q) {
        out[q] = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) out[q] += v[k][q];
    }
    block_sum<N>(out);
}

// The same sum by wavefront 0 alone (no block barrier): lane l adds blocks
// l, l + 64, l + 128, ... in that order (independent loads), then a fixed
// shuffle tree.  Valid in lane 0; called by the 64 lanes of wavefront 0.
template <int N>
__device__ __forceinline__ void wave_sum_partials(const double* __restrict__ part, int nb,
                                                  double (&out)[N]) {
    constexpr int G = 8;   // loads in flight per lane and quantity
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < N; ++q) out[q] = 0.0;
    for (int k0 = 0; k0 * 64 < nb; k0 += G) {
        double v[G][N];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int b = lane + (k0 + k) * 64;
#pragma unroll
            for (int q = 0; q < N; ++q) v[k][q] = b < nb ? part[(size_t)b * N + q] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < G; ++k)
#pragma unroll
            for (int q = 0; q < N; ++q) out[q] += v[k][q];
    }
#pragma unroll
    for (int q = 0; q < N; ++q) out[q] = wave_sum(out[q]);
}

// Sum over a group of L lanes (L <= 64: an xor butterfly inside the wavefront;
// L = 256: the whole block, wavefront sums combined through LDS).  Every lane
// of the group gets the result.  Called with block-uniform control flow.
template <int L, typename Acc> __device__ __forceinline__ Acc group_sum(Acc s) {
    if constexpr (L <= 64) {
        return xor_sum<L>(s);
    } else {
        static_assert(L == kBlock, "block-wide groups span the whole block");
        __shared__ Acc red[kWaves];
        s = xor_sum<64>(s);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
        __syncthreads();
        Acc t = red[0];
#pragma unroll
        for (int k = 1; k < kWaves; ++k) t += red[k];
        __syncthreads();
        return t;
    }
}

// Row dot product with a group of L lanes; every lane of the group gets the
// sum.  valid = false gives 0 but still takes part in the group reduction.
// The CSR loop issues U independent index / value / gather loads before the
// multiply-adds, so a lane has U gathers in flight instead of one; the last
// group of fewer than U is issued the same way under a predicate instead of
// one load pair at a time.  The terms are added in the order p, p + L,
// p + 2L, ... either way.
template <int L, typename Acc, typename V, typename X>
__device__ __forceinline__ Acc row_dot(const MatView<V>& M, const X* __restrict__ x, int row,
                                       int lane, bool valid) {
    constexpr int U = 4;
    Acc s = Acc(0);
    if (M.ell_w > 0) {
        if (valid) {
            const size_t R = (size_t)M.rows;
            auto run = [&](const auto* __restrict__ ix) {
                for (int k = lane; k < M.ell_w; k += L) {
                    const size_t q = (size_t)k * R + row;
                    s += (Acc)M.val[q] * (Acc)x[ix[q]];
                }
            };
            bool done = false;
            if constexpr (std::is_same<V, double>::value) {
                if (M.csc) {
                    const double r = M.rsc[row];
                    auto rebuilt = [&](const auto* __restrict__ ix) {
                        for (int k = lane; k < M.ell_w; k += L) {
                            const int c = ix[(size_t)k * R + row];
                            s += (Acc)(M.csc[c] * r) * (Acc)x[c];
                        }
                    };
                    if (M.idx16) rebuilt(M.idx16);
                    else rebuilt(M.idx);
                    done = true;
                }
            }
            if (!done) {
                if (M.idx16) run(M.idx16);
                else run(M.idx);
            }
        }
    } else {
        int b = 0, e = 0;
        if (valid) { b = M.ptr[row]; e = M.ptr[row + 1]; }
        int p = b + lane;
        for (; p + (U - 1) * L < e; p += U * L) {
            int c[U];
            V v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) { c[u] = M.idx[p + u * L]; v[u] = M.val[p + u * L]; }
            X xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) s += (Acc)v[u] * (Acc)xv[u];
        }
        if (p < e) {
            bool ok[U];
            int c[U];
            V v[U];
            X xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ok[u] = p + u * L < e;
                c[u] = ok[u] ? M.idx[p + u * L] : 0;
                v[u] = ok[u] ? M.val[p + u * L] : V(0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = ok[u] ? x[c[u]] : X(0);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (ok[u]) s += (Acc)v[u] * (Acc)xv[u];
        }
    }
    return group_sum<L, Acc>(s);
}

// Two-element vector of a value type (one 8- or 16-byte access).
template <typename T> struct Pair;
template <> struct Pair<float> { using type = float2; };
template <> struct Pair<double> { using type = double2; };

// Row products of the adjacent ELL rows r and r + 1 (r even, M.rows even) by
// one lane.  The layout is column-major, so the two rows' entries of a slot
// are adjacent: one 4- or 8-byte index access and one 8- or 16-byte value (or
// row factor) access serve both.  Each row adds the same terms in the same
// order as row_dot with one lane per row, so each sum has the same bits.
template <typename T>
__device__ __forceinline__ void ell_pair_dot(const MatView<T>& M, const T* __restrict__ x, int r,
                                             T (&s)[2]) {
    const size_t R = (size_t)M.rows;
    s[0] = T(0);
    s[1] = T(0);
    auto slot = [&](int k, int (&c)[2]) {
        const size_t q = (size_t)k * R + r;
        if (M.idx16) {
            const ushort2 u = *reinterpret_cast<const ushort2*>(M.idx16 + q);
            c[0] = u.x;
            c[1] = u.y;
        } else {
            const int2 u = *reinterpret_cast<const int2*>(M.idx + q);
            c[0] = u.x;
            c[1] = u.y;
        }
    };
    if constexpr (std::is_same<T, double>::value) {
        if (M.csc) {
            const double2 f = *reinterpret_cast<const double2*>(M.rsc + r);
            for (int k = 0; k < M.ell_w; ++k) {
                int c[2];
                slot(k, c);
                s[0] += (M.csc[c[0]] * f.x) * x[c[0]];
                s[1] += (M.csc[c[1]] * f.y) * x[c[1]];
            }
            return;
        }
    }
    for (int k = 0; k < M.ell_w; ++k) {
        int c[2];
        slot(k, c);
        const auto v = *reinterpret_cast<const typename Pair<T>::type*>(M.val + (size_t)k * R + r);
        s[0] += v.x * x[c[0]];
        s[1] += v.y * x[c[1]];
    }
}

// row_dot for long CSR rows with the next group's index and value loads
// issued before this group's gathers are consumed (software pipelining): a
// lane then waits about one memory latency per group of U nonzeros instead
// of two (index, then gather).  Same terms in the same order as row_dot, so
// the same bits; more registers in flight, so it is used where rows are long
// and a kernel has few wavefronts (the dual half).
// row_dot_pipe's loop from entry p on, with the group at p already loaded
// into c, v when it is full (p + (U - 1) L < e), then the predicated last
// group; added into s (the one-pass entry, row_dot_pre).
template <int L, typename Acc, typename V, typename X>
__device__ __forceinline__ Acc csr_pipe_run(const MatView<V>& M, const X* __restrict__ x, int p,
                                            int e, bool full, int (&c)[4], V (&v)[4], Acc s) {
    constexpr int U = 4;
    while (full) {
        X xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
        const int pn = p + U * L;
        const bool nfull = pn + (U - 1) * L < e;
        int cn[U];
        V vn[U];
        if (nfull) {
#pragma unroll
            for (int u = 0; u < U; ++u) { cn[u] = M.idx[pn + u * L]; vn[u] = M.val[pn + u * L]; }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) s += (Acc)v[u] * (Acc)xv[u];
        p = pn;
        full = nfull;
        if (nfull) {
#pragma unroll
            for (int u = 0; u < U; ++u) { c[u] = cn[u]; v[u] = vn[u]; }
        }
    }
    if (p < e) {
        bool ok[U];
        int cc[U];
        V vv[U];
        X xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ok[u] = p + u * L < e;
            cc[u] = ok[u] ? M.idx[p + u * L] : 0;
            vv[u] = ok[u] ? M.val[p + u * L] : V(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = ok[u] ? x[cc[u]] : X(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (ok[u]) s += (Acc)vv[u] * (Acc)xv[u];
    }
    return s;
}

// (row_dot_pipe keeps its own copy of csr_pipe_run's loop: called through
// csr_pipe_run, the grid-stride FP64 dual kernel took 68 instead of 59
// VGPRs, a wavefront per SIMD less.)
template <int L, typename Acc, typename V, typename X>
__device__ __forceinline__ Acc row_dot_pipe(const MatView<V>& M, const X* __restrict__ x, int row,
                                            int lane, bool valid) {
    if (M.ell_w > 0) return row_dot<L, Acc>(M, x, row, lane, valid);
    constexpr int U = 4;
    Acc s = Acc(0);
    int b = 0, e = 0;
    if (valid) { b = M.ptr[row]; e = M.ptr[row + 1]; }
    int p = b + lane;
    bool full = p + (U - 1) * L < e;
    int c[U];
    V v[U];
    if (full) {
#pragma unroll
        for (int u = 0; u < U; ++u) { c[u] = M.idx[p + u * L]; v[u] = M.val[p + u * L]; }
    }
    while (full) {
        X xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
        const int pn = p + U * L;
        const bool nfull = pn + (U - 1) * L < e;
        int cn[U];
        V vn[U];
        if (nfull) {
#pragma unroll
            for (int u = 0; u < U; ++u) { cn[u] = M.idx[pn + u * L]; vn[u] = M.val[pn + u * L]; }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) s += (Acc)v[u] * (Acc)xv[u];
        p = pn;
        full = nfull;
        if (nfull) {
#pragma unroll
            for (int u = 0; u < U; ++u) { c[u] = cn[u]; v[u] = vn[u]; }
        }
    }
    if (p < e) {
        bool ok[U];
        int cc[U];
        V vv[U];
        X xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ok[u] = p + u * L < e;
            cc[u] = ok[u] ? M.idx[p + u * L] : 0;
            vv[u] = ok[u] ? M.val[p + u * L] : V(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = ok[u] ? x[cc[u]] : X(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (ok[u]) s += (Acc)vv[u] * (Acc)xv[u];
    }
    return group_sum<L, Acc>(s);
}

// Two right-hand sides sharing the index stream (FP64, KKT check).
template <int L, typename V>
__device__ __forceinline__ void row_dot2(const MatView<V>& M, const double* __restrict__ x1,
                                         const double* __restrict__ x2, int row, int lane,
                                         bool valid, double& s1, double& s2) {
    constexpr int U = 4;
    s1 = 0.0; s2 = 0.0;
    if (M.ell_w > 0) {
        if (valid) {
            const size_t R = (size_t)M.rows;
            auto run = [&](const auto* __restrict__ ix) {
                for (int k = lane; k < M.ell_w; k += L) {
                    const size_t q = (size_t)k * R + row;
                    const int c = ix[q];
                    const double a = (double)M.val[q];
                    s1 += a * x1[c];
                    s2 += a * x2[c];
                }
            };
            if (M.idx16) run(M.idx16);
            else run(M.idx);
        }
    } else {
        int b = 0, e = 0;
        if (valid) { b = M.ptr[row]; e = M.ptr[row + 1]; }
        int p = b + lane;
        for (; p + (U - 1) * L < e; p += U * L) {
            int c[U];
            double a[U], u1[U], u2[U];
#pragma unroll
            for (int u = 0; u < U; ++u) { c[u] = M.idx[p + u * L]; a[u] = (double)M.val[p + u * L]; }
#pragma unroll
            for (int u = 0; u < U; ++u) { u1[u] = x1[c[u]]; u2[u] = x2[c[u]]; }
#pragma unroll
            for (int u = 0; u < U; ++u) { s1 += a[u] * u1[u]; s2 += a[u] * u2[u]; }
        }
        if (p < e) {
            bool ok[U];
            int c[U];
            double a[U], u1[U], u2[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ok[u] = p + u * L < e;
                c[u] = ok[u] ? M.idx[p + u * L] : 0;
                a[u] = ok[u] ? (double)M.val[p + u * L] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                u1[u] = ok[u] ? x1[c[u]] : 0.0;
                u2[u] = ok[u] ? x2[c[u]] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (ok[u]) { s1 += a[u] * u1[u]; s2 += a[u] * u2[u]; }
        }
    }
    s1 = group_sum<L, double>(s1);
    s2 = group_sum<L, double>(s2);
}

// A xn and A (xn - x) in one pass over the row: one index stream and one
// 8-byte gather of the pair (xn_j, xn_j - x_j) that k_primal stored.  The
// FP32 loop takes the extrapolation A(2 xn - x) as A xn + A (xn - x) and the
// step statistic (yn - y)'A(xn - x) from the difference itself, which keeps
// them accurate when xn and x are close (A xn - A x would cancel to rounding
// noise in FP32).
template <int L>
__device__ __forceinline__ void row_dot_diff(const MatView<float>& M, const float2* __restrict__ xd,
                                             int row, int lane, bool valid, float& s1, float& sd) {
    constexpr int U = 4;
    s1 = 0.f; sd = 0.f;
    if (M.ell_w > 0) {
        if (valid) {
            const size_t R = (size_t)M.rows;
            auto run = [&](const auto* __restrict__ ix) {
                for (int k = lane; k < M.ell_w; k += L) {
                    const size_t q = (size_t)k * R + row;
                    const float a = M.val[q];
                    const float2 v = xd[ix[q]];
                    s1 += a * v.x;
                    sd += a * v.y;
                }
            };
            if (M.idx16) run(M.idx16);
            else run(M.idx);
        }
    } else {
        int b = 0, e = 0;
        if (valid) { b = M.ptr[row]; e = M.ptr[row + 1]; }
        int p = b + lane;
        for (; p + (U - 1) * L < e; p += U * L) {
            int c[U];
            float a[U];
            float2 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) { c[u] = M.idx[p + u * L]; a[u] = M.val[p + u * L]; }
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = xd[c[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) { s1 += a[u] * v[u].x; sd += a[u] * v[u].y; }
        }
        if (p < e) {
            bool ok[U];
            int c[U];
            float a[U];
            float2 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ok[u] = p + u * L < e;
                c[u] = ok[u] ? M.idx[p + u * L] : 0;
                a[u] = ok[u] ? M.val[p + u * L] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = ok[u] ? xd[c[u]] : make_float2(0.f, 0.f);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (ok[u]) { s1 += a[u] * v[u].x; sd += a[u] * v[u].y; }
        }
    }
    s1 = group_sum<L, float>(s1);
    sd = group_sum<L, float>(sd);
}

// ---------------------------------------------- one-pass launches --
//
// When a launch's grid covers every row in a single pass (no grid-stride
// loop, no LDS-staged tiles), a lane's first matrix slots and its row's
// epilogue operands do not depend on the step decision or on the state, so
// the one-pass kernels load them before the decision prologue: their
// latency overlaps the prologue's loads instead of following them.  On a
// small LP an iteration kernel is a chain of such latencies, not bytes.  The
// row products add the same terms in the same order as row_dot /
// row_dot_pipe / row_dot_diff (the prefetched group first, then the rest
// of the row exactly as those loops run it), so the bits are the same.
constexpr int kPreU = 4;
template <typename V> struct PreSlots {
    int next = 0, end = 0;   // CSR: the lane's next entry and the row end; ELL: next slot and width
    int c[kPreU];
    V v[kPreU];
    bool ok[kPreU];
    double r = 0.0;          // ELL values formed from factors: rsc[row]
};

template <int L, typename V>
__device__ __forceinline__ void pre_slots(const MatView<V>& M, int row, int lane, bool valid,
                                          PreSlots<V>& P) {
    if (M.ell_w > 0) {
        const size_t R = (size_t)M.rows;
        bool rebuilt = false;
        if constexpr (std::is_same<V, double>::value) rebuilt = M.csc != nullptr;
#pragma unroll
        for (int u = 0; u < kPreU; ++u) {
            const int k = lane + u * L;
            P.ok[u] = valid && k < M.ell_w;
            const size_t q = (size_t)k * R + row;
            P.c[u] = P.ok[u] ? (M.idx16 ? (int)M.idx16[q] : M.idx[q]) : 0;
            P.v[u] = (P.ok[u] && !rebuilt) ? M.val[q] : V(0);
        }
        if constexpr (std::is_same<V, double>::value) P.r = (rebuilt && valid) ? M.rsc[row] : 0.0;
        P.next = lane + kPreU * L;
        P.end = valid ? M.ell_w : 0;
    } else {
        int b = 0, e = 0;
        if (valid) { b = M.ptr[row]; e = M.ptr[row + 1]; }
        const int p = b + lane;
#pragma unroll
        for (int u = 0; u < kPreU; ++u) {
            P.ok[u] = p + u * L < e;
            P.c[u] = P.ok[u] ? M.idx[p + u * L] : 0;
            P.v[u] = P.ok[u] ? M.val[p + u * L] : V(0);
        }
        P.next = p + kPreU * L;
        P.end = e;
    }
}

// row_dot's CSR loop from entry p on, added into s (groups of 4, then the
// predicated last group).
template <int L, typename Acc, typename V, typename X>
__device__ __forceinline__ Acc csr_dot_from(const MatView<V>& M, const X* __restrict__ x, int p,
                                            int e, Acc s) {
    constexpr int U = 4;
    for (; p + (U - 1) * L < e; p += U * L) {
        int c[U];
        V v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) { c[u] = M.idx[p + u * L]; v[u] = M.val[p + u * L]; }
        X xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) s += (Acc)v[u] * (Acc)xv[u];
    }
    if (p < e) {
        bool ok[U];
        int c[U];
        V v[U];
        X xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ok[u] = p + u * L < e;
            c[u] = ok[u] ? M.idx[p + u * L] : 0;
            v[u] = ok[u] ? M.val[p + u * L] : V(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = ok[u] ? x[c[u]] : X(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (ok[u]) s += (Acc)v[u] * (Acc)xv[u];
    }
    return s;
}

// Long CSR rows in one pass (the dual over few, long rows): a lane's
// entries p, p + L, p + 2L, ... in chunks of kFlatK, each chunk's index and
// value loads issued together and then its gathers together, so a lane waits
// about two memory latencies per chunk; the first chunk is loaded before the
// state.  The terms are added in row_dot_pipe's order, so the same bits.
constexpr int kFlatK = 32;
template <typename V> struct FlatSlots {
    int c[kFlatK];   // -1: past the row's end
    V v[kFlatK];
};

template <int L, typename V>
__device__ __forceinline__ void flat_load(const MatView<V>& M, int p, int e, FlatSlots<V>& F) {
#pragma unroll
    for (int u = 0; u < kFlatK; ++u) {
        const bool ok = p + u * L < e;
        F.c[u] = ok ? M.idx[p + u * L] : -1;
        F.v[u] = ok ? M.val[p + u * L] : V(0);
    }
}

template <typename Acc, typename V, typename X>
__device__ __forceinline__ Acc flat_sum(const FlatSlots<V>& F, const X* __restrict__ x, Acc s) {
    X xv[kFlatK];
#pragma unroll
    for (int u = 0; u < kFlatK; ++u) xv[u] = F.c[u] >= 0 ? x[F.c[u]] : X(0);
#pragma unroll
    for (int u = 0; u < kFlatK; ++u)
        if (F.c[u] >= 0) s += (Acc)F.v[u] * (Acc)xv[u];
    return s;
}

// The row product from the first chunk F (loaded at p) and the rest of the
// row up to e; every lane of the group gets the sum.
template <int L, typename Acc, typename V, typename X>
__device__ __forceinline__ Acc row_dot_flat(const MatView<V>& M, const X* __restrict__ x, int p,
                                            int e, const FlatSlots<V>& F) {
    Acc s = flat_sum<Acc>(F, x, Acc(0));
    for (int q = p + kFlatK * L; q < e; q += kFlatK * L) {
        FlatSlots<V> G;
        flat_load<L>(M, q, e, G);
        s = flat_sum<Acc>(G, x, s);
    }
    return group_sum<L, Acc>(s);
}

// row_dot_diff in the flat form (FP32 loop: A xn and A (xn - x) from the
// interleaved pairs), the terms in row_dot_diff's order.
template <int L>
__device__ __forceinline__ void row_dot_diff_flat(const MatView<float>& M,
                                                  const float2* __restrict__ xd, int p, int e,
                                                  const FlatSlots<float>& F, float& s1, float& sd) {
    s1 = 0.f;
    sd = 0.f;
    auto chunk = [&](const FlatSlots<float>& G) {
        float2 w[kFlatK];
#pragma unroll
        for (int u = 0; u < kFlatK; ++u) w[u] = G.c[u] >= 0 ? xd[G.c[u]] : make_float2(0.f, 0.f);
#pragma unroll
        for (int u = 0; u < kFlatK; ++u)
            if (G.c[u] >= 0) { s1 += G.v[u] * w[u].x; sd += G.v[u] * w[u].y; }
    };
    chunk(F);
    for (int q = p + kFlatK * L; q < e; q += kFlatK * L) {
        FlatSlots<float> G;
        flat_load<L>(M, q, e, G);
        chunk(G);
    }
    s1 = group_sum<L, float>(s1);
    sd = group_sum<L, float>(sd);
}

// The row product from the prefetched slots: row_dot's terms in row_dot's
// order (Pipe: the CSR rest as row_dot_pipe runs it).  Every lane of the
// group gets the sum.
template <int L, bool Pipe, typename Acc, typename V, typename X>
__device__ __forceinline__ Acc row_dot_pre(const MatView<V>& M, const X* __restrict__ x, int row,
                                           const PreSlots<V>& P) {
    Acc s = Acc(0);
    if (M.ell_w > 0) {
        const size_t R = (size_t)M.rows;
        bool rebuilt = false;
        if constexpr (std::is_same<V, double>::value) rebuilt = M.csc != nullptr;
        if (rebuilt) {
            if constexpr (std::is_same<V, double>::value) {
                double f[kPreU];
                X xv[kPreU];
#pragma unroll
                for (int u = 0; u < kPreU; ++u) {
                    f[u] = P.ok[u] ? M.csc[P.c[u]] : 0.0;
                    xv[u] = P.ok[u] ? x[P.c[u]] : X(0);
                }
#pragma unroll
                for (int u = 0; u < kPreU; ++u)
                    if (P.ok[u]) s += (Acc)(f[u] * P.r) * (Acc)xv[u];
                for (int k = P.next; k < P.end; k += L) {
                    const size_t q = (size_t)k * R + row;
                    const int c = M.idx16 ? (int)M.idx16[q] : M.idx[q];
                    s += (Acc)(M.csc[c] * P.r) * (Acc)x[c];
                }
            }
        } else {
            X xv[kPreU];
#pragma unroll
            for (int u = 0; u < kPreU; ++u) xv[u] = P.ok[u] ? x[P.c[u]] : X(0);
#pragma unroll
            for (int u = 0; u < kPreU; ++u)
                if (P.ok[u]) s += (Acc)P.v[u] * (Acc)xv[u];
            for (int k = P.next; k < P.end; k += L) {
                const size_t q = (size_t)k * R + row;
                const int c = M.idx16 ? (int)M.idx16[q] : M.idx[q];
                s += (Acc)M.val[q] * (Acc)x[c];
            }
        }
    } else if constexpr (Pipe) {
        // the prefetched group is row_dot_pipe's first group: when it is full
        // the loop starts from it (its gathers issued beside the next group's
        // index and value loads); otherwise it is the row's last, partial group
        if (P.ok[kPreU - 1]) {
            int c[kPreU];
            V v[kPreU];
#pragma unroll
            for (int u = 0; u < kPreU; ++u) { c[u] = P.c[u]; v[u] = P.v[u]; }
            s = csr_pipe_run<L, Acc>(M, x, P.next - kPreU * L, P.end, true, c, v, s);
        } else {
            X xv[kPreU];
#pragma unroll
            for (int u = 0; u < kPreU; ++u) xv[u] = P.ok[u] ? x[P.c[u]] : X(0);
#pragma unroll
            for (int u = 0; u < kPreU; ++u)
                if (P.ok[u]) s += (Acc)P.v[u] * (Acc)xv[u];
        }
    } else {
        X xv[kPreU];
#pragma unroll
        for (int u = 0; u < kPreU; ++u) xv[u] = P.ok[u] ? x[P.c[u]] : X(0);
#pragma unroll
        for (int u = 0; u < kPreU; ++u)
            if (P.ok[u]) s += (Acc)P.v[u] * (Acc)xv[u];
        s = csr_dot_from<L, Acc>(M, x, P.next, P.end, s);
    }
    return group_sum<L, Acc>(s);
}

// row_dot_diff from the prefetched slots (FP32 loop: A xn and A (xn - x)
// from the interleaved pairs), in row_dot_diff's order.
template <int L>
__device__ __forceinline__ void row_dot_diff_pre(const MatView<float>& M,
                                                 const float2* __restrict__ xd, int row,
                                                 const PreSlots<float>& P, float& s1, float& sd) {
    constexpr int U = 4;
    s1 = 0.f; sd = 0.f;
    float2 v[kPreU];
#pragma unroll
    for (int u = 0; u < kPreU; ++u) v[u] = P.ok[u] ? xd[P.c[u]] : make_float2(0.f, 0.f);
#pragma unroll
    for (int u = 0; u < kPreU; ++u)
        if (P.ok[u]) { s1 += P.v[u] * v[u].x; sd += P.v[u] * v[u].y; }
    if (M.ell_w > 0) {
        const size_t R = (size_t)M.rows;
        for (int k = P.next; k < P.end; k += L) {
            const size_t q = (size_t)k * R + row;
            const float a = M.val[q];
            const float2 w = xd[M.idx16 ? (int)M.idx16[q] : M.idx[q]];
            s1 += a * w.x;
            sd += a * w.y;
        }
    } else {
        int p = P.next;
        const int e = P.end;
        for (; p + (U - 1) * L < e; p += U * L) {
            int c[U];
            float a[U];
            float2 w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) { c[u] = M.idx[p + u * L]; a[u] = M.val[p + u * L]; }
#pragma unroll
            for (int u = 0; u < U; ++u) w[u] = xd[c[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) { s1 += a[u] * w[u].x; sd += a[u] * w[u].y; }
        }
        if (p < e) {
            bool ok[U];
            int c[U];
            float a[U];
            float2 w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ok[u] = p + u * L < e;
                c[u] = ok[u] ? M.idx[p + u * L] : 0;
                a[u] = ok[u] ? M.val[p + u * L] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) w[u] = ok[u] ? xd[c[u]] : make_float2(0.f, 0.f);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (ok[u]) { s1 += a[u] * w[u].x; sd += a[u] * w[u].y; }
        }
    }
    s1 = group_sum<L, float>(s1);
    sd = group_sum<L, float>(sd);
}

// Grid-stride loop over rows, L lanes per row; body(row, lane, valid) runs
// with block-uniform trip counts so group reductions always see full groups.
// XCD-aware mapping: the hardware deals blocks round-robin to the 8 XCDs
// (block b -> XCD b % 8); when the grid is a multiple of 8, block b takes
// the logical slot (b % 8) * (G / 8) + b / 8, so each XCD sweeps one
// contiguous range of rows and the x entries those rows gather stay in that
// XCD's L2.
constexpr int kXcds = 8;
template <int L, typename F>
__device__ __forceinline__ void for_rows(int rows, F&& body) {
    const int gpb = kBlock / L;
    const int g = threadIdx.x / L;
    const int lane = threadIdx.x % L;
    const int G = gridDim.x;
    const int lb = (G % kXcds == 0) ? (blockIdx.x % kXcds) * (G / kXcds) + blockIdx.x / kXcds
                                    : blockIdx.x;
    for (int r0 = lb * gpb; r0 < rows; r0 += G * gpb) {
        const int row = r0 + g;
        body(row, lane, row < rows);
    }
}

// The block's slot after the XCD-aware remap (see for_rows).
__device__ __forceinline__ int xcd_block() {
    const int G = gridDim.x;
    return (G % kXcds == 0) ? (blockIdx.x % kXcds) * (G / kXcds) + blockIdx.x / kXcds
                            : blockIdx.x;
}

// LDS-staged row schedule.  With L lanes per row, for_rows leaves a row's
// epilogue (the projection, averages and stores after the row product) to
// one lane of each group: on a matrix with many rows most of every
// epilogue instruction's lanes are idle and its accesses are scattered.
// Here a block takes tiles of kBlock consecutive rows: it forms the tile's
// row products with L lanes per row (kBlock / L rows per pass, L passes),
// stages the NA values per row in LDS, then thread t runs the epilogue of
// the tile's row t with coalesced accesses.  Used when a matrix has at
// least kStageRows rows (enough tiles to fill the GPU) and L >= 2.
constexpr int kStageRows = 1 << 17;
template <int L> __host__ __device__ __forceinline__ bool staged_rows(int rows) {
    return L >= 2 && rows >= kStageRows;
}

// A launch over `rows` rows with L lanes per row takes one pass when its
// grid (grid_rows: at most kMaxBlocks blocks) holds every row: then each
// thread has one row and lane, for_rows' single iteration.
template <int L> __host__ __device__ __forceinline__ bool one_pass_rows(int rows) {
    return L <= 64 && !staged_rows<L>(rows) && (long)rows <= (long)kMaxBlocks * (kBlock / L);
}

// LDS staging buffer, one per (type, width) whatever the lanes per row, so
// a kernel that dispatches L at run time holds one copy.
template <typename Acc, int NA> __device__ __forceinline__ Acc (&stage_buffer())[NA][kBlock] {
    __shared__ Acc buf[NA][kBlock];
    return buf;
}

// Tiles b0, b0 + nb, b0 + 2 nb, ... of kBlock rows (the single-LP kernels
// pass the remapped block and the grid; a batched instance its own blocks).
template <int L, int NA, typename Acc, typename Dot, typename Epi>
__device__ __forceinline__ void for_row_tiles(int rows, int b0, int nb, Dot&& dot, Epi&& epi) {
    Acc (&stage)[NA][kBlock] = stage_buffer<Acc, NA>();
    constexpr int gpb = L >= kBlock ? 1 : kBlock / L;   // rows per pass
    constexpr int passes = kBlock / gpb;
    const int g = threadIdx.x / (kBlock / gpb);
    const int lane = threadIdx.x % (kBlock / gpb);
    for (int t0 = b0 * kBlock; t0 < rows; t0 += nb * kBlock) {
        for (int pass = 0; pass < passes; ++pass) {
            const int r = pass * gpb + g;
            const int row = t0 + r;
            Acc v[NA];
            dot(row, lane, row < rows, v);
            if (lane == 0) {
#pragma unroll
                for (int q = 0; q < NA; ++q) stage[q][r] = v[q];
            }
        }
        __syncthreads();
        const int row = t0 + threadIdx.x;
        if (row < rows) {
            Acc v[NA];
#pragma unroll
            for (int q = 0; q < NA; ++q) v[q] = stage[q][threadIdx.x];
            epi(row, v);
        }
        __syncthreads();
    }
}

// One row per thread (epilogue-only passes) over blocks b0, b0 + nb, ...
template <typename F> __device__ __forceinline__ void for_rows_flat(int rows, int b0, int nb, F&& body) {
    for (int row = b0 * kBlock + threadIdx.x; row < rows; row += nb * kBlock) body(row);
}

__device__ __forceinline__ double merit3(const double* e) {
    return sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
}

// The step-size rule of include/pdlp_algo.hpp:113-131.
__device__ __forceinline__ void decide_step(DevState* st, double dx2, double dy2,
                                            double inter) {
    st->attempts += 1;
    st->pending = 0;
    const double eta = st->eta, omega = st->omega;
    const double num = omega * dx2 + dy2 / omega;
    const double eta_bar = fabs(inter) > 0 ? num / (2.0 * fabs(inter)) : INFINITY;
    const double k1 = st->accepted + 2.0;
    const bool pre = st->pw_k1 == st->accepted + 2;
    const double p6 = pre ? st->pw6 : pow(k1, -0.6);
    const double p3 = pre ? st->pw3 : pow(k1, -0.3);
    const double grow = (1.0 + p6) * eta;
    const double eta_next = isfinite(eta_bar) ? smin((1.0 - p3) * eta_bar, grow) : grow;
    if (eta <= eta_bar) {
        st->cur ^= 1;
        st->pending = 1;
        st->eta_acc = eta;
        st->wsum += eta;
        st->accepted += 1;
        st->since_restart += 1;
        st->need_aty = 1;
        if (st->accepted % st->check_every == 0) st->need_check = 1;
    } else {
        st->need_aty = 0;
    }
    st->eta = eta_next;
    if (st->attempts >= st->max_iter) st->need_check = 1;
    if (st->need_check) st->active = 0;
}

// Loads the state S[p] into LDS, applies the pending step decision (if the
// previous launch ran an attempt), and publishes the result to S[p ^ 1]
// (the publisher block only).  Returns the block's copy.  Wavefront 0 alone
// loads the state, the primal sum and the dual partials, reduces the
// partials without a block barrier, and decides; one barrier then hands the
// state to the other wavefronts.  The state goes from global memory straight
// into LDS (global_load_lds, one 4-byte word per lane: two wavefront
// instructions for its 74 words), so the copy holds no registers and the
// partials' loads, issued right behind it, are in flight at the same time:
// one memory latency for both.  (Copied through thread 0's registers, the
// state's loads had to land before the partials' were issued, and copying it
// by registers beside the partials' loads took the primal kernels past 100
// VGPRs.)
__device__ __forceinline__ void load_and_decide(DevState* __restrict__ S, int p,
                                                const double* __restrict__ part_m, int nb_m,
                                                const double* __restrict__ dx2_slot,
                                                DevState& Q, bool publisher) {
    if (threadIdx.x < 64) {
        constexpr int W = (int)(sizeof(DevState) / 4);
        static_assert(sizeof(DevState) % 4 == 0 && W > 64 && W <= 128,
                      "the state is copied in two rounds of one word per lane");
        const int lane = (int)threadIdx.x;
        const uint32_t* src = reinterpret_cast<const uint32_t*>(S + p);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&Q);
        // lane l's word lands at dst + l (the LDS base is wavefront-uniform)
        __builtin_amdgcn_global_load_lds(src + lane, dst, 4, 0, 0);
        if (lane < W - 64) __builtin_amdgcn_global_load_lds(src + 64 + lane, dst + 64, 4, 0, 0);
        double dx2 = 0.0;
        if (lane == 0) dx2 = *dx2_slot;
        double sm[2];
        wave_sum_partials<2>(part_m, nb_m, sm);
        // every memory operation of the wavefront, the copy's LDS writes
        // included, has completed before lane 0 reads the state
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        if (threadIdx.x == 0) {
            if (Q.decide_pending) {
                decide_step(&Q, dx2, sm[0], sm[1]);
                Q.decide_pending = 0;
            }
            if (publisher) S[p ^ 1] = Q;
        }
    }
    __syncthreads();
}

// (accepted + 2)^-0.3 and ^-0.6 of the state a launch published to S[p ^ 1],
// stored there for the next decision (one thread of one block).
__device__ __forceinline__ void publish_step_powers(DevState* __restrict__ S, int p, const DevState& Q) {
    const int k1 = Q.accepted + 2;
    DevState* out = S + (p ^ 1);
    out->pw6 = pow((double)k1, -0.6);
    out->pw3 = pow((double)k1, -0.3);
    out->pw_k1 = k1;
}

// KKT check, step 3, for one LP from its summed partials (thread 0):
// eps of both candidates, keep-best, convergence, restart decision
// (include/pdlp_algo.hpp:135-174), and in the mixed mode the request to
// continue in FP64.
__device__ __forceinline__ void kkt_fin_update(DevState* st, const double (&sn)[8],
                                               const double (&sm)[6]) {
    double eps[2][3], obj[2];
    for (int k = 0; k < 2; ++k) {
        const double bv2 = sn[4 * k + 0], svc = sn[4 * k + 1], dcol = sn[4 * k + 2], cx = sn[4 * k + 3];
        const double rv2 = sm[3 * k + 0], svr = sm[3 * k + 1], drow = sm[3 * k + 2];
        const double D = drow + dcol;
        eps[k][0] = sqrt(rv2 + bv2) / st->denom_p;
        eps[k][1] = sqrt(svr + svc) / st->denom_d;
        eps[k][2] = isfinite(D) ? fabs(cx - D) / (1.0 + fabs(cx) + fabs(D)) : INFINITY;
        obj[k] = cx + st->obj_offset;
    }
    st->checks += 1;
    const double mc_cur = merit3(eps[0]), mc_avg = merit3(eps[1]);
    int bw = 0;
    if (mc_cur < st->best_merit) { st->best_merit = mc_cur; bw = 1; }
    if (mc_avg < st->best_merit) { st->best_merit = mc_avg; bw = 2; }
    if (bw) {
        for (int q = 0; q < 3; ++q) st->best_eps[q] = eps[bw - 1][q];
        st->best_obj = obj[bw - 1];
        st->has_best = 1;
    }
    const int use_avg = mc_avg < mc_cur;
    const double* rep = eps[use_avg];
    const double mc = use_avg ? mc_avg : mc_cur;
    for (int q = 0; q < 3; ++q) st->last_eps[q] = rep[q];
    st->last_obj = obj[use_avg];
    st->use_avg = use_avg;
    st->cand_merit = mc;
    st->restart_now = 0;
    if (smax(smax(rep[0], rep[1]), rep[2]) <= st->tol) {
        // converged: the returned point is this candidate
        st->done = 1;
        st->best_which = use_avg ? 2 : 1;
        for (int q = 0; q < 3; ++q) st->best_eps[q] = rep[q];
        st->best_obj = obj[use_avg];
        st->has_best = 1;
        return;
    }
    st->best_which = bw;
    if (st->best_merit < 0.9 * st->stall_ref) { st->stall_ref = st->best_merit; st->stall = 0; }
    else st->stall += 1;
    const bool sufficient = mc <= 0.2 * st->merit_last_restart;
    const bool necessary = mc <= 0.8 * st->merit_last_restart && mc > st->merit_prev_cand;
    const bool artificial = st->since_restart >= 0.36 * st->accepted;
    st->merit_prev_cand = mc;
    if (st->phase == 0 && st->attempts < st->max_iter) {
        // FP32 -> FP64: the merit reached switch_tol, or the best merit stopped
        // improving near the FP32 floor (within 100x switch_tol), or it has
        // stopped improving for four times the stall budget anywhere.
        int why = 0;
        if (mc <= st->switch_tol) why = 1;
        else if (st->stall_checks > 0 && st->stall >= st->stall_checks &&
                 st->best_merit <= 100.0 * st->switch_tol) why = 2;
        else if (st->stall_checks > 0 && st->stall >= 4 * st->stall_checks) why = 3;
        if (why) {
            st->switch_req = why;   // the host restarts this candidate in FP64
            return;
        }
    }
    st->restart_now = (sufficient || necessary || artificial) ? 1 : 0;
    if (st->attempts >= st->max_iter) st->done = 2;
}

// KKT check, step 5, for one LP (thread 0): primal-weight update from the
// squared restart distances dx2, dy2 and the restart counters; resumes the
// iterations unless finished or switching precision.
__device__ __forceinline__ void restart_fin_update(DevState* st, double dx2, double dy2) {
    if (st->restart_now) {
        const double ddx = sqrt(dx2), ddy = sqrt(dy2);
        if (ddx > 1e-10 && ddy > 1e-10)
            st->omega = exp(0.5 * log(ddy / ddx) + 0.5 * log(st->omega));
        st->wsum = 0.0;
        st->since_restart = 0;
        st->restarts += 1;
        st->merit_last_restart = st->cand_merit;
        st->merit_prev_cand = INFINITY;
        if (st->use_avg) st->need_aty = 0;   // A^T y_avg was written with y
    }
    st->pending = 0;
    st->need_check = 0;
    st->best_which = 0;
    st->restart_now = 0;
    if (!st->done && !st->switch_req) st->active = 1;
}

// The FP32 -> FP64 hand-over of the mixed mode, as a restart: the FP64
// iterate starts at the FP32 loop's chosen candidate (copied by the caller,
// with A x and A^T y of it), sums cleared, restart point moved.
__host__ __device__ __forceinline__ void switch_to_fp64(DevState& s) {
    s.attempts_fp32 = s.attempts;
    s.phase = 1;
    s.cur = 0;
    s.switch_req = 0;
    s.wsum = 0.0;
    s.since_restart = 0;
    s.restarts += 1;
    s.merit_last_restart = s.cand_merit;
    s.merit_prev_cand = INFINITY;
    s.pending = 0;
    s.need_check = 0;
    s.need_aty = 0;
    s.decide_pending = 0;
    s.active = 1;
    s.stall = 0;
}


}  // namespace supr
