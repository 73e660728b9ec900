// rule_probe.hip -- TEST INFRASTRUCTURE ONLY: runs the engine's device-side
// scalar rules (pdlp_device.hpp: decide_step, kkt_fin_update,
// restart_fin_update -- the code the single-LP and batched kernels call) on
// scripted inputs, one case per thread, so tests can compare them with the
// oracle's orc_step_rule / orc_check_rule / orc_primal_weight
// (include/pdlp_algo.hpp:94-100, 113-131, 135-167), and the cross-lane
// sums every kernel reduces with against the __shfl forms.  Built into
// tests/hip/librule_probe.so; the product library never loads it.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>
#include <type_traits>

#include "pdlp_device.hpp"

namespace {

using supr::DevState;

__device__ DevState base_state() {
    DevState s;
    memset(&s, 0, sizeof(s));
    s.check_every = 64;
    s.max_iter = INT_MAX;
    s.active = 1;
    s.phase = 1;   // FP64 iterates: no precision-switch requests
    return s;
}

// in (6 per case): eta, omega, accepted, dx2, dy2, inter; pre[t] = 1 seeds the
// step powers the way the previous primal launch publishes them.
// out (2 per case): the next eta, 1.0 if accepted else 0.0.
__global__ void k_step(int n, const double* __restrict__ in, const int* __restrict__ pre,
                       double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double* a = in + 6 * t;
    DevState s = base_state();
    s.eta = a[0];
    s.omega = a[1];
    s.accepted = (int)a[2];
    if (pre[t]) {
        const int k1 = s.accepted + 2;
        s.pw6 = pow((double)k1, -0.6);
        s.pw3 = pow((double)k1, -0.3);
        s.pw_k1 = k1;
    }
    const int acc0 = s.accepted;
    supr::decide_step(&s, a[3], a[4], a[5]);
    out[2 * t] = s.eta;
    out[2 * t + 1] = s.accepted != acc0 ? 1.0 : 0.0;
}

// in (14 + 7 per case): sn[8] (per candidate: bound-violation^2, column
// sign-violation^2, column dual objective, c'x), sm[6] (per candidate: row
// violation^2, row sign-violation^2, row dual objective), then denom_p,
// denom_d, tol, best_merit, merit_last_restart, merit_prev_cand, obj_offset;
// iin (2 per case): since_restart, accepted.
// out (9 per case): done, best_which, use_avg, restart_now, cand_merit,
// best_merit, merit_prev_cand, has_best, checks; eps_out (6 per case): the
// candidate's eps_P, eps_D, eps_G and objective, then best_eps[0..1].
__global__ void k_check(int n, const double* __restrict__ in, const int* __restrict__ iin,
                        double* __restrict__ out, double* __restrict__ eps_out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double* a = in + 21 * t;
    double sn[8], sm[6];
    for (int q = 0; q < 8; ++q) sn[q] = a[q];
    for (int q = 0; q < 6; ++q) sm[q] = a[8 + q];
    DevState s = base_state();
    s.denom_p = a[14];
    s.denom_d = a[15];
    s.tol = a[16];
    s.best_merit = a[17];
    s.merit_last_restart = a[18];
    s.merit_prev_cand = a[19];
    s.obj_offset = a[20];
    s.stall_ref = INFINITY;
    s.since_restart = iin[2 * t];
    s.accepted = iin[2 * t + 1];
    s.attempts = s.accepted;
    s.need_check = 1;
    supr::kkt_fin_update(&s, sn, sm);
    double* o = out + 9 * t;
    o[0] = s.done; o[1] = s.best_which; o[2] = s.use_avg; o[3] = s.restart_now;
    o[4] = s.cand_merit; o[5] = s.best_merit; o[6] = s.merit_prev_cand; o[7] = s.has_best;
    o[8] = s.checks;
    for (int q = 0; q < 3; ++q) eps_out[6 * t + q] = s.last_eps[q];
    eps_out[6 * t + 3] = s.last_obj;
    eps_out[6 * t + 4] = s.best_eps[0];
    eps_out[6 * t + 5] = s.best_eps[1];
}

// in (3 per case): omega, squared restart distances dx2, dy2 (restart due).
// out (1 per case): the primal weight after restart_fin_update.
__global__ void k_weight(int n, const double* __restrict__ in, double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    DevState s = base_state();
    s.omega = in[3 * t];
    s.restart_now = 1;
    s.need_check = 1;
    supr::restart_fin_update(&s, in[3 * t + 1], in[3 * t + 2]);
    out[t] = s.omega;
}

// Cross-lane sums: for each width W = 2..64, the library's DPP / permlane
// butterfly (supr::xor_sum<W>) and the same butterfly through __shfl_xor,
// and supr::wave_sum against the __shfl_down tree it replaced.  One value per
// thread; out holds 6 (W) x 2 (library, shuffle) results per thread, then
// the two wavefront sums.
template <typename T>
__device__ void lane_sums(T v, T* o) {
    int k = 0;
    auto one = [&](auto Wc) {
        constexpr int W = decltype(Wc)::value;
        T ref = v;
#pragma unroll
        for (int q = W / 2; q > 0; q >>= 1) ref += __shfl_xor(ref, q, W);
        o[k++] = supr::xor_sum<W>(v);
        o[k++] = ref;
    };
    one(std::integral_constant<int, 2>{});
    one(std::integral_constant<int, 4>{});
    one(std::integral_constant<int, 8>{});
    one(std::integral_constant<int, 16>{});
    one(std::integral_constant<int, 32>{});
    one(std::integral_constant<int, 64>{});
}

__global__ void k_lanes(int n, const double* __restrict__ in, const float* __restrict__ inf,
                        double* __restrict__ out, float* __restrict__ outf) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;   // n is a multiple of the block
    lane_sums<double>(in[t], out + 14 * (size_t)t);
    lane_sums<float>(inf[t], outf + 12 * (size_t)t);
    double d = in[t];
#pragma unroll
    for (int q = 32; q > 0; q >>= 1) d += __shfl_down(d, q, 64);
    out[14 * (size_t)t + 12] = supr::wave_sum(in[t]);
    out[14 * (size_t)t + 13] = d;
    (void)n;
}

template <typename T> int up(T*& d, const T* h, size_t k) {
    if (hipMalloc((void**)&d, sizeof(T) * (k ? k : 1)) != hipSuccess) return -1;
    if (k && hipMemcpy(d, h, sizeof(T) * k, hipMemcpyHostToDevice) != hipSuccess) return -1;
    return 0;
}

}  // namespace

extern "C" {

// Each entry point copies the cases to the device, runs one thread per
// case, copies the results back and frees everything; 0 on success.
int probe_step_rule(int n, const double* in, const int* pre, double* out) {
    double *din = nullptr, *dout = nullptr;
    int* dpre = nullptr;
    int rc = -1;
    if (n > 0 && !up(din, in, 6 * (size_t)n) && !up(dpre, pre, (size_t)n) &&
        hipMalloc((void**)&dout, sizeof(double) * 2 * n) == hipSuccess) {
        hipLaunchKernelGGL(k_step, dim3((n + 255) / 256), dim3(256), 0, 0, n, din, dpre, dout);
        if (hipGetLastError() == hipSuccess &&
            hipMemcpy(out, dout, sizeof(double) * 2 * n, hipMemcpyDeviceToHost) == hipSuccess)
            rc = 0;
    }
    (void)hipFree(din); (void)hipFree(dpre); (void)hipFree(dout);
    return rc;
}

int probe_check_rule(int n, const double* in, const int* iin, double* out, double* eps_out) {
    double *din = nullptr, *dout = nullptr, *deps = nullptr;
    int* diin = nullptr;
    int rc = -1;
    if (n > 0 && !up(din, in, 21 * (size_t)n) && !up(diin, iin, 2 * (size_t)n) &&
        hipMalloc((void**)&dout, sizeof(double) * 9 * n) == hipSuccess &&
        hipMalloc((void**)&deps, sizeof(double) * 6 * n) == hipSuccess) {
        hipLaunchKernelGGL(k_check, dim3((n + 255) / 256), dim3(256), 0, 0, n, din, diin, dout, deps);
        if (hipGetLastError() == hipSuccess &&
            hipMemcpy(out, dout, sizeof(double) * 9 * n, hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(eps_out, deps, sizeof(double) * 6 * n, hipMemcpyDeviceToHost) == hipSuccess)
            rc = 0;
    }
    (void)hipFree(din); (void)hipFree(diin); (void)hipFree(dout); (void)hipFree(deps);
    return rc;
}

int probe_primal_weight(int n, const double* in, double* out) {
    double *din = nullptr, *dout = nullptr;
    int rc = -1;
    if (n > 0 && !up(din, in, 3 * (size_t)n) && hipMalloc((void**)&dout, sizeof(double) * n) == hipSuccess) {
        hipLaunchKernelGGL(k_weight, dim3((n + 255) / 256), dim3(256), 0, 0, n, din, dout);
        if (hipGetLastError() == hipSuccess &&
            hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost) == hipSuccess)
            rc = 0;
    }
    (void)hipFree(din); (void)hipFree(dout);
    return rc;
}

// n values (a multiple of 256) in in / inf; out: 14 doubles, outf: 12
// floats per value (see k_lanes).
int probe_lane_sums(int n, const double* in, const float* inf, double* out, float* outf) {
    double *din = nullptr, *dout = nullptr;
    float *dinf = nullptr, *doutf = nullptr;
    int rc = -1;
    if (n > 0 && n % 256 == 0 && !up(din, in, (size_t)n) && !up(dinf, inf, (size_t)n) &&
        hipMalloc((void**)&dout, sizeof(double) * 14 * n) == hipSuccess &&
        hipMalloc((void**)&doutf, sizeof(float) * 12 * n) == hipSuccess) {
        hipLaunchKernelGGL(k_lanes, dim3(n / 256), dim3(256), 0, 0, n, din, dinf, dout, doutf);
        if (hipGetLastError() == hipSuccess &&
            hipMemcpy(out, dout, sizeof(double) * 14 * n, hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(outf, doutf, sizeof(float) * 12 * n, hipMemcpyDeviceToHost) == hipSuccess)
            rc = 0;
    }
    (void)hipFree(din); (void)hipFree(dinf); (void)hipFree(dout); (void)hipFree(doutf);
    return rc;
}

}  // extern "C"
