// ranged_lp.hpp -- the LP data model of the PDLP path, with the member names
// of the reference's SparseMatrix / RangedLP (reference: include/sparse.hpp:
// 14-21, include/ranged_lp.hpp:10-25) so code written against those types
// compiles against these:
//   minimize c'x  s.t.  rL <= A x <= rU,  l <= x <= u,  A in COO order.
#pragma once
#include <limits>
#include <string>
#include <vector>

struct SparseMatrix {
    int rows = 0, cols = 0;
    std::vector<int> row_idx;
    std::vector<int> col_idx;
    std::vector<double> val;
    int nnz() const { return static_cast<int>(val.size()); }
};

struct RangedLP {
    std::string name;
    std::vector<std::string> col_names;
    std::vector<std::string> row_names;
    std::vector<double> c;
    SparseMatrix A;
    std::vector<double> rL, rU;
    std::vector<double> l, u;
    double obj_offset = 0.0;
    std::string original_sense = "min";
    int m() const { return static_cast<int>(row_names.size()); }
    int n() const { return static_cast<int>(col_names.size()); }
};

constexpr double INF = std::numeric_limits<double>::infinity();
