// verify_noalloc.hip -- verify() of a device-resident point that allocates
// nothing: supr_verify_gpu_noalloc (include/suprema_pdlp.h).
//
// The LP (CSR of A and of A^T, the vectors) and the point x, y are the
// caller's device memory; the report goes to the caller's device memory too.
// Three kernels run on the caller's stream and the call returns without
// waiting, so it can be captured in a HIP graph or run inside a loop that
// must not allocate:
//   kv_rows  rows of A: (A x)_i, row violations, sign violations of y on
//            infinite row bounds, the rows' part of the dual objective, the
//            finite row-bound norm -> 5 partials per block;
//   kv_cols  rows of A^T (columns): z_j = c_j - (A^T y)_j, bound violations,
//            sign violations of z on infinite column bounds, the columns' part
//            of the dual objective, c'x, the finite |c| norm -> 6 partials;
//   kv_final one block: the partials summed in a fixed order (the same bits
//            every call), then the seven outputs of verify() (src/verify.cpp:
//            24-103) with its norms and finiteness rules.
// The partials live in `scratch`, a fixed device array of this file: both
// passes are grid-stride loops over at most kVerifyBlocks blocks, so its size
// does not depend on the LP.  Each GPU holds its own copy; host code uses the
// address hipGetSymbolAddress gives for it on that device, fetched on the
// first call on each device.  Every call on a GPU uses that GPU's copy:
// calls that overlap on one GPU (two streams at once) race.  Every slot a
// call reads is written by the same call first, so nothing needs clearing.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "hip_util.hpp"
#include "host_lp.hpp"
#include "pdlp_device.hpp"
#include "suprema_pdlp.h"

namespace supr {
namespace {

constexpr int kVerifyBlocks = 1024;            // blocks per pass at most
constexpr int kRowParts = 5, kColParts = 6;
constexpr size_t kScratchDoubles = (size_t)kVerifyBlocks * (kRowParts + kColParts);
// 88 KiB per GPU, held until the process ends
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

// This device's copy of `scratch` (its address, looked up once per device
// with the device current).
double* scratch_on(int device) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch[device] = static_cast<double*>(p);
    }
    return g_scratch[device];
}

template <int L>
__global__ __launch_bounds__(kBlock) void kv_rows(supr_dev_lp lp, const double* __restrict__ x,
                                                  const double* __restrict__ y,
                                                  double* __restrict__ part) {
    MatView<double> A;
    A.rows = lp.m; A.ell_w = 0; A.ptr = lp.A_ptr; A.idx = lp.A_idx; A.val = lp.A_val;
    double sp = 0.0, sd = 0.0, D = 0.0, mx = 0.0, bn = 0.0;
    for_rows<L>(lp.m, [&](int i, int lane, bool valid) {
        const double ax = row_dot<L, double>(A, x, i, lane, valid);
        if (valid && lane == 0) {
            const double lo = lp.rL[i], hi = lp.rU[i];
            const bool lf = isfinite(lo), hf = isfinite(hi);
            const double v = smax(lo - ax, 0.0) + smax(ax - hi, 0.0);
            sp += v * v;
            mx = smax(mx, v);
            const double yi = y[i];
            double s = 0.0;
            if (!lf && yi > 0) s += yi;
            if (!hf && yi < 0) s += -yi;
            sd += s * s;
            double yv = yi;
            if (!lf && yv > 0) yv = 0.0;
            if (!hf && yv < 0) yv = 0.0;
            if (yv > 0 && lf) D += lo * yv;
            if (yv < 0 && hf) D += -hi * -yv;
            if (lf) bn = smax(bn, fabs(lo));
            if (hf) bn = smax(bn, fabs(hi));
        }
    });
    double sums[3] = {sp, sd, D};
    block_sum<3>(sums);
    __shared__ double mxs[kWaves][2];
    double m2[2] = {mx, bn};
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m2[q] = smax(m2[q], __shfl_down(m2[q], o, 64));
    if ((threadIdx.x & 63) == 0) { mxs[threadIdx.x >> 6][0] = m2[0]; mxs[threadIdx.x >> 6][1] = m2[1]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kWaves; ++w) { m2[0] = smax(m2[0], mxs[w][0]); m2[1] = smax(m2[1], mxs[w][1]); }
        double* p = part + (size_t)blockIdx.x * kRowParts;
        p[0] = sums[0]; p[1] = sums[1]; p[2] = sums[2]; p[3] = m2[0]; p[4] = m2[1];
    }
}

template <int L>
__global__ __launch_bounds__(kBlock) void kv_cols(supr_dev_lp lp, const double* __restrict__ x,
                                                  const double* __restrict__ y,
                                                  double* __restrict__ part) {
    MatView<double> AT;
    AT.rows = lp.n; AT.ell_w = 0; AT.ptr = lp.AT_ptr; AT.idx = lp.AT_idx; AT.val = lp.AT_val;
    double sp = 0.0, sd = 0.0, D = 0.0, cx = 0.0, mx = 0.0, cn = 0.0;
    for_rows<L>(lp.n, [&](int j, int lane, bool valid) {
        const double aty = row_dot<L, double>(AT, y, j, lane, valid);
        if (valid && lane == 0) {
            const double lj = lp.l[j], uj = lp.u[j], cj = lp.c[j], xj = x[j];
            const bool lf = isfinite(lj), uf = isfinite(uj);
            const double v = smax(lj - xj, 0.0) + smax(xj - uj, 0.0);
            sp += v * v;
            mx = smax(mx, v);
            const double z = cj - aty;
            double s = 0.0;
            if (!lf && z > 0) s += z;
            if (!uf && z < 0) s += -z;
            sd += s * s;
            double zv = z;
            if (!lf && zv > 0) zv = 0.0;
            if (!uf && zv < 0) zv = 0.0;
            if (zv > 0 && lf) D += lj * zv;
            if (zv < 0 && uf) D += -uj * -zv;
            cx += cj * xj;
            if (isfinite(cj)) cn = smax(cn, fabs(cj));
        }
    });
    double sums[4] = {sp, sd, D, cx};
    block_sum<4>(sums);
    __shared__ double mxs[kWaves][2];
    double m2[2] = {mx, cn};
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m2[q] = smax(m2[q], __shfl_down(m2[q], o, 64));
    if ((threadIdx.x & 63) == 0) { mxs[threadIdx.x >> 6][0] = m2[0]; mxs[threadIdx.x >> 6][1] = m2[1]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kWaves; ++w) { m2[0] = smax(m2[0], mxs[w][0]); m2[1] = smax(m2[1], mxs[w][1]); }
        double* p = part + (size_t)blockIdx.x * kColParts;
        p[0] = sums[0]; p[1] = sums[1]; p[2] = sums[2]; p[3] = sums[3]; p[4] = m2[0]; p[5] = m2[1];
    }
}

// One block: thread 0 adds the partials in block order (the same bits every
// call) and writes verify()'s seven outputs.
__global__ __launch_bounds__(kBlock) void kv_final(const double* __restrict__ prow, int gm,
                                                   const double* __restrict__ pcol, int gn, int m,
                                                   int n, double obj_offset,
                                                   double* __restrict__ out7) {
    if (threadIdx.x != 0) return;
    double sp = 0.0, sd = 0.0, Dr = 0.0, mrv = 0.0, bn = 0.0;
    for (int b = 0; b < gm; ++b) {
        const double* p = prow + (size_t)b * kRowParts;
        sp += p[0]; sd += p[1]; Dr += p[2]; mrv = smax(mrv, p[3]); bn = smax(bn, p[4]);
    }
    double Dc = 0.0, cx = 0.0, mbv = 0.0, cn = 0.0;
    for (int b = 0; b < gn; ++b) {
        const double* p = pcol + (size_t)b * kColParts;
        sp += p[0]; sd += p[1]; Dc += p[2]; cx += p[3]; mbv = smax(mbv, p[4]); cn = smax(cn, p[5]);
    }
    const double D = Dr + Dc;
    out7[0] = sqrt(sp) / (1.0 + bn);
    out7[1] = sqrt(sd) / (1.0 + cn);
    out7[2] = isfinite(D) ? fabs(cx - D) / (1.0 + fabs(cx) + fabs(D)) : INFINITY;
    out7[3] = cx + obj_offset;
    out7[4] = isfinite(D) ? D + obj_offset : D;
    out7[5] = m > 0 ? mrv : 0.0;
    out7[6] = n > 0 ? mbv : 0.0;
}

// Lanes per row from the mean row length (any choice gives the same report
// within rounding; this one keeps long rows off a single lane).
int lanes_for_mean(int64_t nnz, int rows) {
    const double mean = rows > 0 ? (double)nnz / rows : 0.0;
    return mean <= 4.0 ? 1 : (mean <= 64.0 ? 8 : 32);
}

int grid_for(int rows, int L) {
    const long per = kBlock / L;
    long g = ((long)rows + per - 1) / per;
    if (g < 1) g = 1;
    if (g > kVerifyBlocks) g = kVerifyBlocks;
    if (g >= kXcds) g = g / kXcds * kXcds;   // XCD-aware mapping needs a multiple of 8
    return (int)g;
}

}  // namespace
}  // namespace supr

extern "C" int supr_verify_gpu_noalloc(const supr_dev_lp* lp, const double* x, const double* y,
                                       double* out7, void* stream) {
    using namespace supr;
    if (!lp || !out7 || (lp->n > 0 && !x) || (lp->m > 0 && !y)) {
        set_error("supr_verify_gpu_noalloc: null argument");
        return SUPR_EINVAL;
    }
    if (lp->m < 0 || lp->n < 0 || lp->nnz < 0) {
        set_error("supr_verify_gpu_noalloc: negative sizes");
        return SUPR_EINVAL;
    }
    if ((lp->m > 0 && (!lp->A_ptr || !lp->rL || !lp->rU)) || (lp->n > 0 && (!lp->AT_ptr || !lp->c || !lp->l || !lp->u)) ||
        (lp->nnz > 0 && (!lp->A_idx || !lp->A_val || !lp->AT_idx || !lp->AT_val))) {
        set_error("supr_verify_gpu_noalloc: null LP array");
        return SUPR_EINVAL;
    }
    try {
        int dev = 0;
        SUPR_HIP(hipGetDevice(&dev));
        if (dev < 0 || dev >= kMaxDevices) {
            set_error("supr_verify_gpu_noalloc: device ordinal beyond the library's table");
            return SUPR_EINVAL;
        }
        double* sc = scratch_on(dev);
        double* prow = sc;
        double* pcol = sc + (size_t)kVerifyBlocks * kRowParts;
        hipStream_t s = static_cast<hipStream_t>(stream);
        const int la = lanes_for_mean(lp->nnz, lp->m), lat = lanes_for_mean(lp->nnz, lp->n);
        const int gm = grid_for(lp->m, la), gn = grid_for(lp->n, lat);
        if (gm > kVerifyBlocks || gn > kVerifyBlocks) {   // the scratch holds kVerifyBlocks per pass
            set_error("supr_verify_gpu_noalloc: more blocks than the scratch array holds");
            return SUPR_EINTERNAL;
        }
        switch (la) {
            case 1: hipLaunchKernelGGL(kv_rows<1>, dim3(gm), dim3(kBlock), 0, s, *lp, x, y, prow); break;
            case 8: hipLaunchKernelGGL(kv_rows<8>, dim3(gm), dim3(kBlock), 0, s, *lp, x, y, prow); break;
            default: hipLaunchKernelGGL(kv_rows<32>, dim3(gm), dim3(kBlock), 0, s, *lp, x, y, prow); break;
        }
        switch (lat) {
            case 1: hipLaunchKernelGGL(kv_cols<1>, dim3(gn), dim3(kBlock), 0, s, *lp, x, y, pcol); break;
            case 8: hipLaunchKernelGGL(kv_cols<8>, dim3(gn), dim3(kBlock), 0, s, *lp, x, y, pcol); break;
            default: hipLaunchKernelGGL(kv_cols<32>, dim3(gn), dim3(kBlock), 0, s, *lp, x, y, pcol); break;
        }
        hipLaunchKernelGGL(kv_final, dim3(1), dim3(kBlock), 0, s, prow, gm, pcol, gn, lp->m, lp->n,
                           lp->obj_offset, out7);
        SUPR_HIP(hipGetLastError());
        return SUPR_OK;
    } catch (...) {
        return map_exception();
    }
}
