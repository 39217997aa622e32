// extern "C" entry points over the reference's own C++ CSR spmv, so the test
// suite can call it through ctypes.  Compiled together with
// qutip/core/data/src/matmul_csr_vector.cpp, where it lies in the reference
// checkout (recipe: oracle/Makefile); the output goes to oracle/_ref/ only.
// The wrapper adds nothing to the arithmetic: it forwards to the reference's
// _matmul_csr_vector<int> (matmul_csr_vector.cpp:107-149 portable path, as
// built with no -m flags like the reference's meson build).
#include <complex>
#include "matmul_csr_vector.hpp"

extern "C" void qtref_csr_matvec(const double *data, const int *col_index,
                                 const int *row_index, const double *vec,
                                 double scale_re, double scale_im, double *out,
                                 int nrows)
{
    _matmul_csr_vector<int>(reinterpret_cast<const std::complex<double> *>(data),
                            col_index, row_index,
                            reinterpret_cast<const std::complex<double> *>(vec),
                            std::complex<double>(scale_re, scale_im),
                            reinterpret_cast<std::complex<double> *>(out), nrows);
}
