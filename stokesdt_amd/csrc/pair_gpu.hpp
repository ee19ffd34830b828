// Periodic cell-list pair lists on the GPU (PairListR / PairListS).
#pragma once
#include "sdt_internal.hpp"
#include "sort_gpu.hpp"

namespace sdt {

struct PairListGpu {
    // configuration (src/pairlist/pair_base.h:30-69)
    int npos = 0;
    double box = 0, cutoff = 0;
    char mode = 'R';
    double max_radius = 0, min_radius = 0;
    int nc1 = 0;
    double cell_size = 0;
    int nsidx = 0;
    std::vector<int> h_rdi_dummy;
    // device
    DevBuf<double> d_rdi;       // radii, original order
    DevBuf<int> d_sidx;         // 4 ints per stencil entry: dx, dy, dz, all
    DevBuf<double> d_ssum;      // 2 doubles per stencil entry: min_sum, max_sum
    DevBuf<double> d_pos0;      // wrapped positions
    DevBuf<int> d_cidx;         // cell of each particle
    DevBuf<int> d_cell_perm, d_cell_start;
    DevBuf<int> d_cnt, d_rowptr, d_colidx;
    DevBuf<int> d_nnz;          // 1 int (host-readable copy in h_nnz)
    DevBuf<int> d_dcidx, d_dperm, d_dstart;   // contact-distance cells (steric_direct)
    SortScratch sc;
    int nnz = 0;
    long long colcap = 0;

    // returns the estimated nnz (> 0), 0 on invalid input (PairListBase::Init)
    int init(int npos, const double* h_rdi, double box, double cutoff, char mode);
    // Builds the list for device positions. rows/cols in original order
    // (row_map/col_map null) or remapped (row_map[r] = particle of row r,
    // col_map[j] = column id of particle j); sort_rows sorts each row
    // ascending. Returns nnz (synchronises once to size the column array).
    int build(const double* d_pos, const int* row_map, const int* col_map, bool sort_rows,
              hipStream_t s);
    // steric mode: f += the steric forces of the pairs build() lists, from
    // cells as wide as the contact distance (no CSR, no host
    // synchronisation); equal to build + steric_accumulate up to the order
    // of each particle's FP64 sum
    void steric_direct(const double* d_pos, double r0, double k0, double* d_f, hipStream_t s);
};

// Steric forces (src/force/force_steric.cc:14-53) on a pair list in
// original order: f_i += k0 ((a_i + a_j) r0 / 2 - r) (r_i - r_j) / r.
void steric_accumulate(int npos, const double* d_pos, const double* d_rdi, double box, double r0,
                       double k0, const int* d_rowptr, const int* d_colidx, double* d_f,
                       hipStream_t s);

}  // namespace sdt
