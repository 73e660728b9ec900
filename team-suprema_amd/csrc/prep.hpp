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
    // Uploads lp (validated by the caller); with_layout builds A, AT, vA, vAT.
    DevLp(const supr_lp& lp, hipStream_t s, bool with_layout = true);
    // Takes over device arrays of an LP with the given sizes and lays it out.
    DevLp(int32_t m, int32_t n, int32_t nnz, DevBuf&& row, DevBuf&& col, DevBuf&& val, DevBuf&& c,
          DevBuf&& l, DevBuf&& u, DevBuf&& rL, DevBuf&& rU, hipStream_t s);
    int32_t m = 0, n = 0, nnz = 0;
    DevBuf row, col, val, c, l, u, rL, rU;
    DevPattern A, AT;
    DevBuf vA, vAT;   // FP64 values in CSR (A) and CSC (A^T) order
    void layout(hipStream_t s);
};

// build_pdlp_form (src/pdlp.cpp:8-82) of a DevLp, on the device: the form's
// COO arrays and vectors, and the meta row_perm / row_signs (m2 entries).
struct DevForm {
    int32_t m2 = 0, n2 = 0, nnz2 = 0, n_slack = 0, n_orig = 0, m_orig = 0;
    DevBuf row, col, val, c, l, u, rL, rU, row_perm, row_signs;
};
void build_form_on_device(const DevLp& lp, DevForm& f, hipStream_t s);

// ax = A x and aty = A^T y (device vectors) over a laid-out DevLp, each row
// summed left to right in COO order like the host verifier's matvecs.
void verify_products_on_device(const DevLp& lp, const double* x, const double* y, double* ax,
                               double* aty, hipStream_t s);

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

// COO-order values of the pattern's LP in the pattern's order (CSR for A,
// CSC for A^T): out = a new device array of P.nnz values.
void gather_values(const DevPattern& P, const double* coo_val, DevBuf& out, hipStream_t s);

}  // namespace supr
