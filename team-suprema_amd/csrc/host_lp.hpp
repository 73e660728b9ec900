// host_lp.hpp -- host-side LP helpers shared by the engine and the C ABI:
// CSR construction, the FP64 KKT verifier, scaling, the PDLP form, error
// reporting.  Internal header (not installed).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "suprema_pdlp.h"

namespace supr {

// Thread-local error text for supr_last_error().
void set_error(const std::string& msg);

// Row-major CSR.  Entries of a row keep their COO order (stable counting sort),
// the layout of the reference's to_csr (include/sparse.hpp:31-47), so row sums
// visit terms in the same order.
struct Csr {
    int32_t rows = 0, cols = 0;
    std::vector<int32_t> ptr;   // rows + 1
    std::vector<int32_t> idx;   // nnz
    std::vector<double> val;    // nnz
    // perm[k] = COO position of CSR entry k, so other value arrays with the
    // same pattern can be laid out identically.
    std::vector<int32_t> perm;
};

// transposed = false: CSR of A (rows = m).  true: CSR of A^T (rows = n), the
// reference's to_csc_as_transposed_csr (include/sparse.hpp:50-57).
Csr build_csr(const supr_lp& lp, const double* values, bool transposed);

// out[i] = sum_p val[p] * x[idx[p]], left to right.
void csr_matvec(const Csr& M, const double* x, double* out);

// Validates shapes and index ranges; sets the error text and returns false
// when the LP is malformed.
bool check_lp(const supr_lp* lp, const char* what, bool indices = true);

// verify() with prebuilt CSR(A) / CSR(A^T) of lp's values (src/verify.cpp:24-103).
void verify_csr(const supr_lp& lp, const Csr& A, const Csr& AT, const double* x,
                const double* y, double* out7);

// The same from the products ax = A x and aty = A^T y (left-to-right row
// sums of the CSR / CSC): the O(m + n) tail of verify(), src/verify.cpp:34-103.
void verify_products(const supr_lp& lp, const double* x, const double* y, const double* ax,
                     const double* aty, double* out7);

}  // namespace supr
