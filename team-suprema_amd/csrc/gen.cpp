// gen.cpp -- seeded synthetic LP generators behind supr_gen_transport and
// supr_gen_random_lp.  The PRNG is this file's own (splitmix64-seeded
// xoshiro256**, fixed transforms), so an LP is a pure function of its
// arguments on every platform.
//
// Draw order (part of the generators' definition):
//   transport: S supplies, D raw demands, then S*D costs in (i, j) order.
//   random LP: per column j (in order) its distinct rows then its values;
//              then per-column upper bounds (only if u_hi < 0), x0, y0, s0.
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "host_lp.hpp"

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& w : s) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ULL;
            t = (t ^ (t >> 27)) * 0x94d049bb133111ebULL;
            w = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // [0, 1) with 53 random bits
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    // integer in [a, b] (multiply-shift)
    int64_t integer(int64_t a, int64_t b) {
        const uint64_t range = (uint64_t)(b - a) + 1;
        return a + (int64_t)(((unsigned __int128)next() * range) >> 64);
    }
    // N(0,1), Box-Muller (one variate per call)
    double normal() {
        const double u1 = 1.0 - uniform();   // (0, 1]
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

}  // namespace

extern "C" {

int supr_gen_transport(int32_t S, int32_t D, uint64_t seed, int32_t cost_max, int32_t* row_idx,
                       int32_t* col_idx, double* val, double* c, double* rL, double* rU,
                       double* l, double* u) {
    if (S < 2 || D < 2 || cost_max < 2 || (int64_t)S * D > (int64_t)INT32_MAX / 2) {
        supr::set_error("supr_gen_transport: need S, D >= 2, cost_max >= 2, 2*S*D < 2^31");
        return SUPR_EINVAL;
    }
    Rng rng(seed);
    std::vector<int64_t> supply(S), raw(D), demand(D);
    int64_t total = 0, rawsum = 0;
    for (int32_t i = 0; i < S; ++i) { supply[i] = rng.integer(50, 150); total += supply[i]; }
    for (int32_t j = 0; j < D; ++j) { raw[j] = rng.integer(50, 150); rawsum += raw[j]; }
    int64_t assigned = 0;
    for (int32_t j = 0; j < D - 1; ++j) {
        int64_t dj = (int64_t)((double)total * (double)raw[j] / (double)rawsum);
        if (dj < 1) dj = 1;
        demand[j] = dj;
        assigned += dj;
    }
    demand[D - 1] = total - assigned;
    if (demand[D - 1] < 1) {
        supr::set_error("supr_gen_transport: infeasible demand split for this seed");
        return SUPR_EINVAL;
    }
    const double inf = INFINITY;
    int64_t k = 0;
    for (int32_t i = 0; i < S; ++i) {
        for (int32_t j = 0; j < D; ++j) {
            const int32_t col = i * D + j;
            c[col] = (double)rng.integer(1, cost_max);
            l[col] = 0.0;
            u[col] = inf;
            row_idx[k] = i; col_idx[k] = col; val[k] = 1.0; ++k;
            row_idx[k] = S + j; col_idx[k] = col; val[k] = 1.0; ++k;
        }
    }
    for (int32_t i = 0; i < S; ++i) rL[i] = rU[i] = (double)supply[i];
    for (int32_t j = 0; j < D; ++j) rL[S + j] = rU[S + j] = (double)demand[j];
    return SUPR_OK;
}

int supr_gen_random_lp(int32_t m, int32_t n, int32_t col_nnz, const int32_t* col_deg,
                       uint64_t seed, double x0_hi, double u_hi, int32_t mag_decades,
                       int32_t* row_idx, int32_t* col_idx, double* val, double* c, double* rL,
                       double* rU, double* l, double* u) {
    if (m < 1 || n < 1 || (!col_deg && (col_nnz < 1 || col_nnz > m)) || x0_hi < 0) {
        supr::set_error("supr_gen_random_lp: bad sizes");
        return SUPR_EINVAL;
    }
    int64_t nnz = 0;
    for (int32_t j = 0; j < n; ++j) {
        const int32_t d = col_deg ? col_deg[j] : col_nnz;
        if (d < 1 || d > m) { supr::set_error("supr_gen_random_lp: column degree out of [1, m]"); return SUPR_EINVAL; }
        nnz += d;
    }
    if (nnz > INT32_MAX) { supr::set_error("supr_gen_random_lp: nnz exceeds int32"); return SUPR_EINVAL; }
    Rng rng(seed);
    // Structure and values, column by column.  Distinct rows by rejection
    // against the column's rows drawn so far (degrees are << m).
    std::vector<int32_t> rows;
    int64_t k = 0;
    for (int32_t j = 0; j < n; ++j) {
        const int32_t d = col_deg ? col_deg[j] : col_nnz;
        rows.clear();
        while ((int32_t)rows.size() < d) {
            const int32_t r = (int32_t)rng.integer(0, m - 1);
            bool dup = false;
            for (int32_t q : rows) dup |= (q == r);
            if (!dup) rows.push_back(r);
        }
        for (int32_t r : rows) {
            double v;
            if (mag_decades > 0) {
                const double mag = std::pow(10.0, rng.uniform(-(double)mag_decades, (double)mag_decades));
                v = (rng.next() >> 63) ? -mag : mag;
            } else {
                v = rng.normal();
            }
            row_idx[k] = r; col_idx[k] = j; val[k] = v; ++k;
        }
    }
    // Bounds: [0, u_hi], or [0, U[1, -u_hi]] per column when u_hi < 0.
    for (int32_t j = 0; j < n; ++j) {
        l[j] = 0.0;
        u[j] = u_hi < 0 ? rng.uniform(1.0, -u_hi) : u_hi;
    }
    // Feasible point x0 inside the box; b = A x0 accumulated in COO order.
    std::vector<double> x0(n), y0(m);
    for (int32_t j = 0; j < n; ++j) x0[j] = rng.uniform(0.0, std::fmin(x0_hi, u[j]));
    for (int32_t i = 0; i < m; ++i) rL[i] = 0.0;
    for (int64_t q = 0; q < nnz; ++q) rL[row_idx[q]] += val[q] * x0[col_idx[q]];
    for (int32_t i = 0; i < m; ++i) rU[i] = rL[i];
    // Bounded: c = A'y0 + s0 with s0 >= 0 (the box bounds it anyway).
    for (int32_t i = 0; i < m; ++i) y0[i] = rng.normal();
    for (int32_t j = 0; j < n; ++j) c[j] = 0.0;
    for (int64_t q = 0; q < nnz; ++q) c[col_idx[q]] += val[q] * y0[row_idx[q]];
    for (int32_t j = 0; j < n; ++j) c[j] += rng.uniform();
    return SUPR_OK;
}

}  // extern "C"

extern "C" {

// Blending LP (see suprema_pdlp.h): K components, P products, two quality
// attributes.  Columns x_{kp} in product-major order (col = p*K + k).  Rows:
// P demand equalities, K capacity rows (<=), 2P ranged quality rows.
// Draw order: a_{kq} (K x 2), d_p (P), blend weights w_{kp} (P x K), costs (P x K).
int supr_gen_blend(int32_t K, int32_t P, uint64_t seed, int32_t* row_idx, int32_t* col_idx,
                   double* val, double* c, double* rL, double* rU, double* l, double* u) {
    if (K < 1 || P < 1 || (int64_t)K * P * 4 > INT32_MAX) {
        supr::set_error("supr_gen_blend: need K, P >= 1 and 4*K*P < 2^31");
        return SUPR_EINVAL;
    }
    Rng rng(seed);
    const int32_t m = 3 * P + K;
    std::vector<double> a((size_t)K * 2), d(P), w((size_t)P * K), cap(K, 0.0), qual((size_t)P * 2, 0.0);
    for (auto& v : a) v = rng.uniform(0.5, 1.5);
    for (auto& v : d) v = rng.uniform(50.0, 150.0);
    for (int32_t p = 0; p < P; ++p) {
        double s = 0.0;
        for (int32_t k = 0; k < K; ++k) { w[(size_t)p * K + k] = rng.uniform(); s += w[(size_t)p * K + k]; }
        for (int32_t k = 0; k < K; ++k) {
            const double wk = w[(size_t)p * K + k] / s;          // a feasible blend of product p
            w[(size_t)p * K + k] = wk;
            cap[k] += wk * d[p];
            qual[(size_t)p * 2] += a[(size_t)k * 2] * wk;
            qual[(size_t)p * 2 + 1] += a[(size_t)k * 2 + 1] * wk;
        }
    }
    const double inf = INFINITY;
    int64_t q = 0;
    for (int32_t p = 0; p < P; ++p) {
        for (int32_t k = 0; k < K; ++k) {
            const int32_t col = p * K + k;
            c[col] = rng.uniform(10.0, 100.0);
            l[col] = 0.0;
            u[col] = inf;
            row_idx[q] = p; col_idx[q] = col; val[q] = 1.0; ++q;                    // demand
            row_idx[q] = P + k; col_idx[q] = col; val[q] = 1.0; ++q;                // capacity
            row_idx[q] = P + K + p; col_idx[q] = col; val[q] = a[(size_t)k * 2]; ++q;          // quality 0
            row_idx[q] = P + K + P + p; col_idx[q] = col; val[q] = a[(size_t)k * 2 + 1]; ++q;  // quality 1
        }
    }
    for (int32_t p = 0; p < P; ++p) { rL[p] = rU[p] = d[p]; }
    for (int32_t k = 0; k < K; ++k) { rL[P + k] = -inf; rU[P + k] = 1.1 * cap[k]; }
    for (int32_t p = 0; p < P; ++p)
        for (int32_t t = 0; t < 2; ++t) {
            const int32_t r = P + K + t * P + p;
            rL[r] = 0.9 * qual[(size_t)p * 2 + t] * d[p];
            rU[r] = 1.1 * qual[(size_t)p * 2 + t] * d[p];
        }
    (void)m;
    return SUPR_OK;
}

// Per-instance sizes of a blending batch: K log-uniform in [K_lo, K_hi], P
// log-uniform in [P_lo, P_hi] (rounded), from this file's PRNG.
int supr_gen_blend_sizes(int32_t count, uint64_t seed, int32_t K_lo, int32_t K_hi, int32_t P_lo,
                         int32_t P_hi, int32_t* K, int32_t* P) {
    if (count < 0 || K_lo < 1 || K_hi < K_lo || P_lo < 1 || P_hi < P_lo) {
        supr::set_error("supr_gen_blend_sizes: bad ranges");
        return SUPR_EINVAL;
    }
    Rng rng(seed);
    for (int32_t i = 0; i < count; ++i) {
        K[i] = (int32_t)std::lround(std::exp(rng.uniform(std::log((double)K_lo), std::log((double)K_hi))));
        P[i] = (int32_t)std::lround(std::exp(rng.uniform(std::log((double)P_lo), std::log((double)P_hi))));
    }
    return SUPR_OK;
}

}  // extern "C"
