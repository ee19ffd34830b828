// Matrix-free real-space RPY-Ewald matvec.
#pragma once
#include "sdt_internal.hpp"

namespace sdt {

// Y = beta Y + alpha * M_real X over a CSR pair list (self included) whose
// rows and columns index the same particle order as pos/rdi and the packed
// vectors X, Y (row-major 3N x b).
template <typename T>
void real_matvec(int npos, const int* d_rowptr, const int* d_colidx, const double* d_pos,
                 const double* d_rdi, double box, double xi, int b, double alpha, const double* X,
                 double beta, double* Y, hipStream_t s);

}  // namespace sdt
