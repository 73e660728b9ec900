// prep.hpp -- the LP's matrix laid out and scaled on the device: the
// reference's to_csr / to_csc_as_transposed_csr (include/sparse.hpp:31-57)
// and scale() (src/scaling.cpp:25-84), with results bit-identical to the
// host versions.  Internal header.
#pragma once
#include <cstdint>

#include "hip_util.hpp"

namespace supr {

// CSR of a COO matrix built on the device: entries of each row keep their
// COO order; perm[p] is the COO position of entry p.
struct DevPattern {
    int32_t rows = 0, cols = 0, nnz = 0;
    int32_t min_len = 0, max_len = 0;   // shortest and longest row
    DevBuf ptr, idx, perm;              // rows + 1, nnz, nnz
};

// An LP on the device: the COO arrays and vectors as given, the patterns of
// A (rows) and A^T (columns), and A's values in both orders.
class DevLp {
   public:
    DevLp(const supr_lp& lp, hipStream_t s);
    int32_t m = 0, n = 0, nnz = 0;
    DevBuf row, col, val, c, l, u, rL, rU;
    DevPattern A, AT;
    DevBuf vA, vAT;   // FP64 values in CSR (A) and CSC (A^T) order
};

// scale() of a DevLp on the device.
struct DevScaled {
    DevBuf vA, vAT;                   // scaled values, CSR / CSC order
    DevBuf Dr, Dc, c, l, u, rL, rU;   // factors and scaled vectors
};

// integer: host flags (n bytes) of columns kept unscaled, or null.
void scale_on_device(const DevLp& lp, const char* integer, int ruiz_passes, DevScaled& out,
                     hipStream_t s);

// max |v| over len device values (eta0 = 1 / max|a_ij|).
double device_absmax(const double* v, int32_t len, hipStream_t s);

// out[perm[p]] = v[p] for the pattern's entries (device arrays).
void scatter_to_coo(const DevPattern& P, const double* v, double* out, hipStream_t s);

}  // namespace supr
