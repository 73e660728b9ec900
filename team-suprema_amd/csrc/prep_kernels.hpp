// prep_kernels.hpp -- device preprocessing in front of the PDHG loop: the
// stable COO -> CSR / CSC build and the reference's scale() (src/scaling.cpp:
// 25-84: 10 Ruiz passes, then Pock-Chambolle), computed so that every value
// has the bits the host computation gives.
//
// Bit-exactness rests on three facts.
//  * Maxima are order-free, so the Ruiz row / column maxima may be reduced
//    in any order (the reference scans COO order, scaling.cpp:4-13).
//  * The Pock-Chambolle sums are sequential in COO order (scaling.cpp:14-23).
//    The CSR (CSC) built here keeps COO order inside every row (column): the
//    sort below is a stable LSD radix sort of COO positions by row (column)
//    index, like the reference's counting sort (include/sparse.hpp:31-57).
//    Each row's sum is then one left-to-right chain of FP64 adds: one thread
//    per short row, one wavefront per long row (coalesced loads, the chain
//    runs over the 64 loaded terms in lane order).
//  * Every scaled value is the same product chain as the reference's:
//    per Ruiz pass a <- a * r_i, then a <- a * c_j (the reference's
//    multiplications by 1.0 in scale_rows_cols, include/sparse.hpp:124-129,
//    change no bits), and finally a <- (a * Dr_pc_i) * Dc_pc_j.  The
//    translation unit is built with floating-point contraction off, and
//    1/sqrt is the correctly rounded division of the correctly rounded
//    square root, as in the host.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace supr {
namespace prep {

constexpr int kThreads = 256;              // 4 wavefronts per block
constexpr int kItems = 16;                 // items per thread per tile
constexpr int kTile = kThreads * kItems;   // COO entries per sort tile
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kLongRow = 256;   // rows longer than this sum with a whole wavefront
constexpr int kGrid = 2048;     // grid-stride cap for row and element loops

__device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }

// ------------------------------------------------------------------ sort --

// Per-tile digit histogram, digit-major: hist[d * tiles + tile].
__global__ __launch_bounds__(kThreads) void kp_radix_hist(const int32_t* __restrict__ keys,
                                                          int32_t n, int shift,
                                                          int32_t* __restrict__ hist,
                                                          int32_t tiles) {
    __shared__ int32_t cnt[kRadix];
    const int tid = threadIdx.x;
    const int32_t tile = blockIdx.x;
    cnt[tid] = 0;
    __syncthreads();
    for (int r = 0; r < kItems; ++r) {
        const int64_t i = (int64_t)tile * kTile + r * kThreads + tid;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & (kRadix - 1)], 1);
    }
    __syncthreads();
    hist[(int64_t)tid * tiles + tile] = cnt[tid];
}

// Stable scatter of one radix pass.  offs is the exclusive scan of the
// digit-major histogram, so offs[d * tiles + tile] is where this tile's
// first entry with digit d goes.  Entries are ranked in position order:
// round r of the tile holds positions r*256 + tid, and an entry's rank is
// (entries of its digit in earlier rounds) + (in earlier wavefronts of its
// round) + (in lower lanes of its wavefront).  vals_in == nullptr means the
// values are the positions themselves (the first pass).
__global__ __launch_bounds__(kThreads) void kp_radix_scatter(
    const int32_t* __restrict__ keys_in, const int32_t* __restrict__ vals_in, int32_t n, int shift,
    const int32_t* __restrict__ offs, int32_t tiles, int32_t* __restrict__ keys_out,
    int32_t* __restrict__ vals_out) {
    __shared__ int32_t base[kRadix];
    __shared__ int32_t wcnt[kThreads / 64][kRadix];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int32_t tile = blockIdx.x;
    base[tid] = offs[(int64_t)tid * tiles + tile];
    for (int k = 0; k < kThreads / 64; ++k) wcnt[k][tid] = 0;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int r = 0; r < kItems; ++r) {
        const int64_t i = (int64_t)tile * kTile + r * kThreads + tid;
        const bool valid = i < n;
        int32_t key = 0, val = 0;
        int d = 0;
        if (valid) {
            key = keys_in[i];
            val = vals_in ? vals_in[i] : (int32_t)i;
            d = (key >> shift) & (kRadix - 1);
        }
        // lanes of this wavefront holding the same digit
        uint64_t same = __ballot(valid);
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1;
            const uint64_t set = __ballot(valid && bit);
            same &= bit ? set : ~set;
        }
        __syncthreads();   // the previous round's reads of base / wcnt are done
        if (r > 0) {
            int32_t t = 0;
            for (int k = 0; k < kThreads / 64; ++k) { t += wcnt[k][tid]; wcnt[k][tid] = 0; }
            base[tid] += t;
        }
        __syncthreads();
        if (valid && (same & below) == 0) wcnt[w][d] = __popcll(same);
        __syncthreads();
        if (valid) {
            int32_t pos = base[d] + __popcll(same & below);
            for (int k = 0; k < w; ++k) pos += wcnt[k][d];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
    }
}

// ------------------------------------------------------------------ scan --

// Exclusive scan of kThreads values held one per thread; *total gets the sum.
__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t* total) {
    __shared__ int32_t wsum[kThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int32_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int32_t off = 0, tot = 0;
    for (int k = 0; k < kThreads / 64; ++k) {
        off += k < w ? wsum[k] : 0;
        tot += wsum[k];
    }
    __syncthreads();   // wsum may be reused by the caller's next call
    *total = tot;
    return off + x - v;
}

// Tile-local exclusive scan (thread t owns items t*16 .. t*16+15 of the
// tile); totals[tile] = the tile's sum.
__global__ __launch_bounds__(kThreads) void kp_scan_tiles(const int32_t* in,
                                                          int32_t* out, int64_t len,
                                                          int32_t* __restrict__ totals) {
    const int tid = threadIdx.x;
    const int64_t b0 = (int64_t)blockIdx.x * kTile + (int64_t)tid * kItems;
    int32_t v[kItems];
    int32_t run = 0;
    for (int k = 0; k < kItems; ++k) {
        v[k] = (b0 + k < len) ? in[b0 + k] : 0;
        const int32_t t = v[k];
        v[k] = run;
        run += t;
    }
    int32_t tot;
    const int32_t off = block_exclusive_scan(run, &tot);
    for (int k = 0; k < kItems; ++k)
        if (b0 + k < len) out[b0 + k] = v[k] + off;
    if (tid == 0) totals[blockIdx.x] = tot;
}

// In-place exclusive scan of the tile totals by one block.
__global__ __launch_bounds__(kThreads) void kp_scan_totals(int32_t* __restrict__ totals,
                                                           int32_t count) {
    int32_t carry = 0;
    for (int32_t b0 = 0; b0 < count; b0 += kThreads) {
        const int32_t i = b0 + threadIdx.x;
        const int32_t v = i < count ? totals[i] : 0;
        int32_t tot;
        const int32_t ex = block_exclusive_scan(v, &tot);
        if (i < count) totals[i] = carry + ex;
        carry += tot;
    }
}

__global__ __launch_bounds__(kThreads) void kp_scan_add(int32_t* __restrict__ out, int64_t len,
                                                        const int32_t* __restrict__ totals) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * kThreads)
        out[i] += totals[i / kTile];
}

// --------------------------------------------------------------- pattern --

// Row pointers from the sorted keys: entry p starts every row in
// (key[p-1], key[p]], and the last entry closes the rows after it.
__global__ __launch_bounds__(kThreads) void kp_ptr_from_sorted(const int32_t* __restrict__ key,
                                                               int32_t nnz, int32_t rows,
                                                               int32_t* __restrict__ ptr) {
    for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p <= nnz;
         p += (int64_t)gridDim.x * kThreads) {
        const int32_t prev = p > 0 ? key[p - 1] : -1;
        const int32_t cur = p < nnz ? key[p] : rows;
        for (int32_t i = prev + 1; i <= cur; ++i) ptr[i] = (int32_t)p;
    }
}

// idx[p] = other[perm[p]].
__global__ __launch_bounds__(kThreads) void kp_gather_idx(const int32_t* __restrict__ perm,
                                                          int32_t nnz,
                                                          const int32_t* __restrict__ other,
                                                          int32_t* __restrict__ idx) {
    for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < nnz;
         p += (int64_t)gridDim.x * kThreads)
        idx[p] = other[perm[p]];
}

// out[p] = coo_val[perm[p]].
__global__ __launch_bounds__(kThreads) void kp_gather_val(const int32_t* __restrict__ perm,
                                                          int32_t nnz,
                                                          const double* __restrict__ coo_val,
                                                          double* __restrict__ out) {
    for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < nnz;
         p += (int64_t)gridDim.x * kThreads)
        out[p] = coo_val[perm[p]];
}

// out[perm[p]] = v[p]: values back in COO order.
__global__ __launch_bounds__(kThreads) void kp_scatter_coo(const int32_t* __restrict__ perm,
                                                           int32_t nnz, const double* __restrict__ v,
                                                           double* __restrict__ out) {
    for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < nnz;
         p += (int64_t)gridDim.x * kThreads)
        out[perm[p]] = v[p];
}

// ------------------------------------------------------------- scaling --

__global__ __launch_bounds__(kThreads) void kp_fill(double* __restrict__ v, int32_t len, double x) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * kThreads)
        v[i] = x;
}

// One Ruiz half-pass over the rows of a CSR (src/scaling.cpp:36-51):
// a <- a * g[idx] first when g is given (the other side's factor of the
// previous half-pass, not yet applied to this copy), then
// s = max|a| > 0 and not kept ? 1/sqrt(max|a|) : 1, a <- a * s, D *= s,
// sout = s.  L lanes per row; the maximum is order-free.
template <int L>
__global__ __launch_bounds__(kThreads) void kp_ruiz(int32_t rows, const int32_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ idx,
                                                    double* __restrict__ v,
                                                    const double* __restrict__ g,
                                                    const char* __restrict__ keep,
                                                    double* __restrict__ D,
                                                    double* __restrict__ sout) {
    constexpr int G = kThreads / L;
    const int grp = threadIdx.x / L, sub = threadIdx.x % L;
    for (int64_t row = (int64_t)blockIdx.x * G + grp; row < rows; row += (int64_t)gridDim.x * G) {
        const int32_t b = ptr[row], e = ptr[row + 1];
        double mx = 0.0;
        for (int32_t p = b + sub; p < e; p += L) {
            double a = v[p];
            if (g) a = a * g[idx[p]];
            mx = dmax(mx, fabs(a));
        }
        for (int o = L / 2; o > 0; o >>= 1) mx = dmax(mx, __shfl_xor(mx, o, L));
        const bool kept = keep != nullptr && keep[row] != 0;
        const double s = (mx > 0.0 && !kept) ? 1.0 / sqrt(mx) : 1.0;
        for (int32_t p = b + sub; p < e; p += L) {
            double a = v[p];
            if (g) a = a * g[idx[p]];
            v[p] = a * s;
        }
        if (sub == 0) {
            D[row] = D[row] * s;
            sout[row] = s;
        }
    }
}

// Pock-Chambolle factor of a row from its abs-sum (src/scaling.cpp:55-60):
// sqrt(sum > 0 ? 1/sum : 1), or 1 where the row is kept.
__device__ __forceinline__ double pc_factor(double sum, bool kept) {
    return kept ? 1.0 : sqrt(sum > 0.0 ? 1.0 / sum : 1.0);
}

// Sequential abs-sums of rows with at most kLongRow entries, one thread per
// row: a <- a * g[idx] (stored) when g is given, then sum |a| left to right.
__global__ __launch_bounds__(kThreads) void kp_pc_short(int32_t rows,
                                                        const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ idx,
                                                        double* __restrict__ v,
                                                        const double* __restrict__ g,
                                                        const char* __restrict__ keep,
                                                        double* __restrict__ S) {
    for (int64_t row = (int64_t)blockIdx.x * kThreads + threadIdx.x; row < rows;
         row += (int64_t)gridDim.x * kThreads) {
        const int32_t b = ptr[row], e = ptr[row + 1];
        if (e - b > kLongRow) continue;
        double sum = 0.0;
        for (int32_t p = b; p < e; ++p) {
            double a = v[p];
            if (g) {
                a = a * g[idx[p]];
                v[p] = a;
            }
            sum = sum + fabs(a);
        }
        S[row] = pc_factor(sum, keep != nullptr && keep[row] != 0);
    }
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// The same for rows longer than kLongRow, one wavefront per row: each lane
// loads one term of a 64-term chunk, then the chain adds the chunk's terms
// in lane order (every lane holds the same running sum).
__global__ __launch_bounds__(kThreads) void kp_pc_long(int32_t rows,
                                                       const int32_t* __restrict__ ptr,
                                                       const int32_t* __restrict__ idx,
                                                       double* __restrict__ v,
                                                       const double* __restrict__ g,
                                                       const char* __restrict__ keep,
                                                       double* __restrict__ S) {
    const int lane = threadIdx.x & 63;
    constexpr int W = kThreads / 64;
    for (int64_t row = (int64_t)blockIdx.x * W + (threadIdx.x >> 6); row < rows;
         row += (int64_t)gridDim.x * W) {
        // one row per wavefront: its bounds are uniform (scalar registers)
        const int32_t b = __builtin_amdgcn_readfirstlane(ptr[row]);
        const int32_t e = __builtin_amdgcn_readfirstlane(ptr[row + 1]);
        if (e - b <= kLongRow) continue;
        double sum = 0.0;
        for (int32_t c0 = b; c0 < e; c0 += 64) {
            const int32_t p = c0 + lane;
            double a = 0.0;
            if (p < e) {
                a = v[p];
                if (g) {
                    a = a * g[idx[p]];
                    v[p] = a;
                }
            }
            a = fabs(a);
            const int cnt = e - c0 < 64 ? e - c0 : 64;
            for (int t = 0; t < cnt; ++t) sum = sum + readlane_f64(a, t);
        }
        if (lane == 0) S[row] = pc_factor(sum, keep != nullptr && keep[row] != 0);
    }
}

// Final Pock-Chambolle application (src/scaling.cpp:63): a <- (a * R) * C
// where R is Dr_pc of the entry's row and C is Dc_pc of its column.  For a
// CSR of A, R is indexed by the row and C by idx; for the CSC, the reverse.
template <int L>
__global__ __launch_bounds__(kThreads) void kp_pc_apply(int32_t rows,
                                                        const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ idx,
                                                        double* __restrict__ v,
                                                        const double* __restrict__ f_row,
                                                        const double* __restrict__ f_idx,
                                                        int row_first) {
    constexpr int G = kThreads / L;
    const int grp = threadIdx.x / L, sub = threadIdx.x % L;
    for (int64_t row = (int64_t)blockIdx.x * G + grp; row < rows; row += (int64_t)gridDim.x * G) {
        const int32_t b = ptr[row], e = ptr[row + 1];
        const double fr = f_row[row];
        for (int32_t p = b + sub; p < e; p += L) {
            const double fi = f_idx[idx[p]];
            v[p] = row_first ? (v[p] * fr) * fi : (v[p] * fi) * fr;
        }
    }
}

// D *= F elementwise.
__global__ __launch_bounds__(kThreads) void kp_mul(double* __restrict__ D,
                                                   const double* __restrict__ F, int32_t len) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * kThreads)
        D[i] = D[i] * F[i];
}

// Scaled vectors (src/scaling.cpp:68-76): c~ = Dc c, l~ = l / Dc, u~ = u / Dc.
__global__ __launch_bounds__(kThreads) void kp_scale_cols(int32_t n, const double* __restrict__ Dc,
                                                          const double* __restrict__ c,
                                                          const double* __restrict__ l,
                                                          const double* __restrict__ u,
                                                          double* __restrict__ cs,
                                                          double* __restrict__ ls,
                                                          double* __restrict__ us) {
    for (int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * kThreads) {
        const double d = Dc[j];
        cs[j] = d * c[j];
        ls[j] = l[j] / d;
        us[j] = u[j] / d;
    }
}

// r~L = Dr rL, r~U = Dr rU.
__global__ __launch_bounds__(kThreads) void kp_scale_rows(int32_t m, const double* __restrict__ Dr,
                                                          const double* __restrict__ rL,
                                                          const double* __restrict__ rU,
                                                          double* __restrict__ rLs,
                                                          double* __restrict__ rUs) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * kThreads) {
        const double d = Dr[i];
        rLs[i] = d * rL[i];
        rUs[i] = d * rU[i];
    }
}

// Per-block max |v| (for eta0 = 1 / max|a_ij|, include/pdlp_algo.hpp:70).
__global__ __launch_bounds__(kThreads) void kp_absmax(const double* __restrict__ v, int32_t len,
                                                      double* __restrict__ part) {
    __shared__ double sm[kThreads / 64];
    double mx = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * kThreads)
        mx = dmax(mx, fabs(v[i]));
    for (int o = 32; o > 0; o >>= 1) mx = dmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kThreads / 64; ++k) mx = dmax(mx, sm[k]);
        part[blockIdx.x] = mx;
    }
}

// ---------------------------------------------------- verifier products --

// out[row] = sum_p v[p] * x[idx[p]] left to right, each product rounded and
// then added, as the host's sequential matvec (include/sparse.hpp:59-66):
// one thread per row of at most kLongRow entries ...
__global__ __launch_bounds__(kThreads) void kp_spmv_seq_short(int32_t rows,
                                                              const int32_t* __restrict__ ptr,
                                                              const int32_t* __restrict__ idx,
                                                              const double* __restrict__ v,
                                                              const double* __restrict__ x,
                                                              double* __restrict__ out) {
    for (int64_t row = (int64_t)blockIdx.x * kThreads + threadIdx.x; row < rows;
         row += (int64_t)gridDim.x * kThreads) {
        const int32_t b = ptr[row], e = ptr[row + 1];
        if (e - b > kLongRow) continue;
        double sum = 0.0;
        for (int32_t p = b; p < e; ++p) sum = sum + v[p] * x[idx[p]];
        out[row] = sum;
    }
}

// ... and one wavefront per longer row (each lane forms one product of a
// 64-term chunk; the chain adds them in lane order).
__global__ __launch_bounds__(kThreads) void kp_spmv_seq_long(int32_t rows,
                                                             const int32_t* __restrict__ ptr,
                                                             const int32_t* __restrict__ idx,
                                                             const double* __restrict__ v,
                                                             const double* __restrict__ x,
                                                             double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    constexpr int W = kThreads / 64;
    for (int64_t row = (int64_t)blockIdx.x * W + (threadIdx.x >> 6); row < rows;
         row += (int64_t)gridDim.x * W) {
        const int32_t b = __builtin_amdgcn_readfirstlane(ptr[row]);
        const int32_t e = __builtin_amdgcn_readfirstlane(ptr[row + 1]);
        if (e - b <= kLongRow) continue;
        double sum = 0.0;
        for (int32_t c0 = b; c0 < e; c0 += 64) {
            const int32_t p = c0 + lane;
            const double t = p < e ? v[p] * x[idx[p]] : 0.0;
            const int cnt = e - c0 < 64 ? e - c0 : 64;
            for (int k = 0; k < cnt; ++k) sum = sum + readlane_f64(t, k);
        }
        if (lane == 0) out[row] = sum;
    }
}

// z = c - aty (src/verify.cpp:33).
__global__ __launch_bounds__(kThreads) void kp_sub(int32_t len, const double* __restrict__ c,
                                                   const double* __restrict__ a,
                                                   double* __restrict__ z) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < len;
         i += (int64_t)gridDim.x * kThreads)
        z[i] = c[i] - a[i];
}

// ------------------------------------------------------------ PDLP form --
// build_pdlp_form (src/pdlp.cpp:8-82): equality and ranged rows first, in
// their original order; then <= rows (negated into >=) and >= rows; free rows
// dropped; a ranged row becomes an equality row with b = 0 and a slack
// column (coefficient -1) carrying the range.

enum FormRowKind : int { kRowEq = 0, kRowRanged, kRowLe, kRowGe, kRowFree };

__device__ __forceinline__ int form_row_kind(double lo, double hi) {
    if (lo == hi) return kRowEq;
    if (isfinite(lo) && isfinite(hi)) return kRowRanged;
    if (isfinite(hi)) return kRowLe;
    if (isfinite(lo)) return kRowGe;
    return kRowFree;
}

// fE: equality or ranged; fI: <= or >=; fR: ranged.
__global__ __launch_bounds__(kThreads) void kf_flags(int32_t m, const double* __restrict__ rL,
                                                     const double* __restrict__ rU,
                                                     int32_t* __restrict__ fE,
                                                     int32_t* __restrict__ fI,
                                                     int32_t* __restrict__ fR) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * kThreads) {
        const int k = form_row_kind(rL[i], rU[i]);
        fE[i] = (k == kRowEq || k == kRowRanged);
        fI[i] = (k == kRowLe || k == kRowGe);
        fR[i] = (k == kRowRanged);
    }
}

// totals[t] = scan[t][len-1] + flag[t][len-1] for the three scans (len > 0).
__global__ void kf_totals(int32_t len, const int32_t* __restrict__ sE,
                          const int32_t* __restrict__ fE, const int32_t* __restrict__ sI,
                          const int32_t* __restrict__ fI, const int32_t* __restrict__ sR,
                          const int32_t* __restrict__ fR, int32_t* __restrict__ totals) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        totals[0] = sE[len - 1] + fE[len - 1];
        totals[1] = sI[len - 1] + fI[len - 1];
        totals[2] = sR[len - 1] + fR[len - 1];
    }
}

// Row placement: row_perm, row_signs, rL2, rU2 (m2 entries), pos_of (m),
// slack_pos (the form row of slack s) and the slack bounds l2, u2[n + s].
__global__ __launch_bounds__(kThreads) void kf_rows(
    int32_t m, int32_t n, int32_t n_eq, const double* __restrict__ rL,
    const double* __restrict__ rU, const int32_t* __restrict__ posE,
    const int32_t* __restrict__ posI, const int32_t* __restrict__ sid,
    int32_t* __restrict__ row_perm, double* __restrict__ row_signs, double* __restrict__ rL2,
    double* __restrict__ rU2, int32_t* __restrict__ pos_of, int32_t* __restrict__ slack_pos,
    double* __restrict__ l2, double* __restrict__ u2) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < m;
         i += (int64_t)gridDim.x * kThreads) {
        const double lo = rL[i], hi = rU[i];
        const int k = form_row_kind(lo, hi);
        if (k == kRowFree) { pos_of[i] = -1; continue; }
        const int32_t p = (k == kRowEq || k == kRowRanged) ? posE[i] : n_eq + posI[i];
        pos_of[i] = p;
        row_perm[p] = (int32_t)i;
        row_signs[p] = k == kRowLe ? -1.0 : 1.0;
        if (k == kRowEq) {
            rL2[p] = lo;
            rU2[p] = lo;
        } else if (k == kRowRanged) {
            rL2[p] = 0.0;
            rU2[p] = 0.0;
            const int32_t sl = sid[i];
            slack_pos[sl] = p;
            l2[n + sl] = lo;
            u2[n + sl] = hi;
        } else {
            rL2[p] = k == kRowLe ? -hi : lo;
            rU2[p] = INFINITY;
        }
    }
}

// fK[k] = the entry's row is kept.
__global__ __launch_bounds__(kThreads) void kf_keep(int32_t nnz, const int32_t* __restrict__ row,
                                                    const int32_t* __restrict__ pos_of,
                                                    int32_t* __restrict__ fK) {
    for (int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x; k < nnz;
         k += (int64_t)gridDim.x * kThreads)
        fK[k] = pos_of[row[k]] >= 0;
}

// Kept entries in COO order (q = exclusive scan of fK), then the slacks'.
__global__ __launch_bounds__(kThreads) void kf_entries(
    int32_t nnz, const int32_t* __restrict__ row, const int32_t* __restrict__ col,
    const double* __restrict__ val, const int32_t* __restrict__ pos_of,
    const double* __restrict__ row_signs, const int32_t* __restrict__ fK,
    const int32_t* __restrict__ q, int32_t* __restrict__ row2, int32_t* __restrict__ col2,
    double* __restrict__ val2) {
    for (int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x; k < nnz;
         k += (int64_t)gridDim.x * kThreads) {
        if (!fK[k]) continue;
        const int32_t p = pos_of[row[k]], Q = q[k];
        row2[Q] = p;
        col2[Q] = col[k];
        val2[Q] = row_signs[p] * val[k];
    }
}

__global__ __launch_bounds__(kThreads) void kf_slack_entries(int32_t n_slack, int32_t base,
                                                             int32_t n,
                                                             const int32_t* __restrict__ slack_pos,
                                                             int32_t* __restrict__ row2,
                                                             int32_t* __restrict__ col2,
                                                             double* __restrict__ val2) {
    for (int64_t sl = (int64_t)blockIdx.x * kThreads + threadIdx.x; sl < n_slack;
         sl += (int64_t)gridDim.x * kThreads) {
        row2[base + sl] = slack_pos[sl];
        col2[base + sl] = n + (int32_t)sl;
        val2[base + sl] = -1.0;
    }
}

}  // namespace prep
}  // namespace supr
