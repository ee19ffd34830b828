// Oracle pair lists (test infrastructure only; see orc_common.hpp).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <omp.h>
#include "orc_pairlist.hpp"

namespace orc {

PairList::PairList(int npos, const double* rdi, double box, double cutoff, char mode)
    : npos_(npos), box_(box), cutoff_(cutoff), mode_(mode),
      rdi_(rdi, rdi + (npos > 0 ? npos : 0)) {}

// src/pairlist/pair_r.cc:15-22 and pair_s.cc:16-24
int PairList::num_cells() const {
    int nc1;
    if (mode_ == 'R') nc1 = (int)(box_ / (cutoff_ / 2.0));
    else nc1 = (int)(box_ / (cutoff_ * max_radius_ / 2.0));
    return nc1 == 0 ? 1 : nc1;
}

// src/pairlist/pair_r.cc:25-68 and pair_s.cc:27-71: stencil of cell
// offsets, with the FP64 distance bounds that decide which offsets are
// searched and which are taken whole.
void PairList::init_search_idx() {
    const int nc1 = nc1_;
    const size_t nc2 = (size_t)nc1 * nc1;
    const size_t nc3 = nc2 * nc1;
    sidx_.clear();
    for (size_t i = 0; i < nc3; i++) {
        int ix = (int)(i / nc2);
        int iy = (int)((i % nc2) / nc1);
        int iz = (int)(i % nc1);
        ix = ix > nc1 / 2 ? ix - nc1 : ix;
        iy = iy > nc1 / 2 ? iy - nc1 : iy;
        iz = iz > nc1 / 2 ? iz - nc1 : iz;
        double dx = abs(ix) - 1 > 0 ? abs(ix) - 1 : 0;
        double dy = abs(iy) - 1 > 0 ? abs(iy) - 1 : 0;
        double dz = abs(iz) - 1 > 0 ? abs(iz) - 1 : 0;
        double rmin = sqrt(dx * dx + dy * dy + dz * dz) * cell_size_;
        dx = abs(ix) + 1 > nc1 / 2 ? abs(ix) + 0.5 : abs(ix) + 1;
        dy = abs(iy) + 1 > nc1 / 2 ? abs(iy) + 0.5 : abs(iy) + 1;
        dz = abs(iz) + 1 > nc1 / 2 ? abs(iz) + 0.5 : abs(iz) + 1;
        double rmax = sqrt(dx * dx + dy * dy + dz * dz) * cell_size_;
        Sidx s;
        s.dx = ix; s.dy = iy; s.dz = iz;
        if (mode_ == 'R') {
            if (rmin > cutoff_) continue;
            s.all = (rmax <= cutoff_);
            s.min_sum = s.max_sum = 0.0;
        } else {
            double min_sum = 2.0 * rmin / cutoff_;
            if (min_sum > 2.0 * max_radius_) continue;
            s.all = false;
            s.min_sum = min_sum;
            s.max_sum = 2.0 * rmax / cutoff_;
        }
        sidx_.push_back(s);
    }
}

// estimated nnz (src/pairlist/pair_r.cc:71-92, pair_s.cc:74-96)
int PairList::init_pairs() {
    int nnz = 0;
    pairs_.assign(npos_, std::vector<int>());
    for (int i = 0; i < npos_; i++) {
        double t1;
        if (mode_ == 'R') t1 = pow(cutoff_ + max_radius_, 3.0);
        else t1 = pow((rdi_[i] + max_radius_) * cutoff_ / 2.0 + max_radius_, 3.0);
        double t2 = pow(rdi_[i], 3.0);
        double t3 = pow(min_radius_, 3.0);
        int size = (int)((t1 - t2) / t3 * 0.6) + 1;
        size = size > npos_ ? npos_ : size;
        pairs_[i].reserve(size);
        nnz += size;
    }
    return nnz;
}

// src/pairlist/pair_base.cc:34-71
int PairList::init() {
    if (npos_ <= 0 || box_ <= 0.0 || cutoff_ <= 0.0) return 0;
    if (mode_ != 'R' && mode_ != 'S') return 0;
    max_radius_ = *std::max_element(rdi_.begin(), rdi_.end());
    min_radius_ = *std::min_element(rdi_.begin(), rdi_.end());
    if (min_radius_ <= 0.0) return 0;
    nc1_ = num_cells();
    cell_size_ = box_ / nc1_;
    size_t nc3 = (size_t)nc1_ * nc1_ * nc1_;
    head_.assign(nc3, -1);
    next_.assign(npos_, -1);
    cidx_.assign(npos_, 0);
    pos0_.assign((size_t)npos_ * 3, 0.0);
    init_search_idx();
    return init_pairs();
}

// src/pairlist/pair_base.cc:74-115 (serial linked-list build; the
// reference's per-thread ownership only changes the list order, and rows
// are sorted afterwards)
void PairList::construct_cells(const double* pos) {
    const size_t nc1 = nc1_;
    const size_t nc2 = nc1 * nc1;
    std::fill(head_.begin(), head_.end(), -1);
    for (int i = 0; i < npos_; i++) {
        double x = fmod(pos[3 * i + 0], box_);
        double y = fmod(pos[3 * i + 1], box_);
        double z = fmod(pos[3 * i + 2], box_);
        x = (x >= 0.0 ? x : box_ + x);
        y = (y >= 0.0 ? y : box_ + y);
        z = (z >= 0.0 ? z : box_ + z);
        int ix = (int)(x / box_ * nc1_);
        int iy = (int)(y / box_ * nc1_);
        int iz = (int)(z / box_ * nc1_);
        size_t idx = ix * nc2 + iy * nc1 + iz;
        cidx_[i] = idx;
        pos0_[3 * i + 0] = x;
        pos0_[3 * i + 1] = y;
        pos0_[3 * i + 2] = z;
        next_[i] = head_[idx];
        head_[idx] = i;
    }
}

static inline int wrap_cell(int c, int nc1) {
    c = c >= nc1 ? c - nc1 : c;
    return c < 0 ? c + nc1 : c;
}

// src/pairlist/pair_r.cc:95-170
int PairList::find_pairs_r() {
    int nnz = 0;
    const int nc1 = nc1_;
    const size_t nc2 = (size_t)nc1 * nc1;
    const int ns = (int)sidx_.size();
    #pragma omp parallel for schedule(dynamic, 64) reduction(+ : nnz)
    for (int i = 0; i < npos_; i++) {
        size_t idx1 = cidx_[i];
        double x1 = pos0_[3 * i], y1 = pos0_[3 * i + 1], z1 = pos0_[3 * i + 2];
        int ix1 = (int)(idx1 / nc2), iy1 = (int)((idx1 % nc2) / nc1), iz1 = (int)(idx1 % nc1);
        std::vector<int>& row = pairs_[i];
        row.clear();
        for (int k = 0; k < ns; k++) {
            int ix2 = wrap_cell(ix1 + sidx_[k].dx, nc1);
            int iy2 = wrap_cell(iy1 + sidx_[k].dy, nc1);
            int iz2 = wrap_cell(iz1 + sidx_[k].dz, nc1);
            size_t idx2 = ix2 * nc2 + iy2 * nc1 + iz2;
            int j = head_[idx2];
            if (sidx_[k].all) {
                for (; j != -1; j = next_[j]) row.push_back(j);
            } else {
                for (; j != -1; j = next_[j]) {
                    double rx = fabs(pos0_[3 * j] - x1);
                    double ry = fabs(pos0_[3 * j + 1] - y1);
                    double rz = fabs(pos0_[3 * j + 2] - z1);
                    rx = (rx > box_ / 2.0 ? rx - box_ : rx);
                    ry = (ry > box_ / 2.0 ? ry - box_ : ry);
                    rz = (rz > box_ / 2.0 ? rz - box_ : rz);
                    double rr = rx * rx + ry * ry + rz * rz;
                    if (sqrt(rr) <= cutoff_) row.push_back(j);
                }
            }
        }
        std::sort(row.begin(), row.end());
        nnz += (int)row.size();
    }
    return nnz;
}

// src/pairlist/pair_s.cc:99-186
int PairList::find_pairs_s() {
    int nnz = 0;
    const int nc1 = nc1_;
    const size_t nc2 = (size_t)nc1 * nc1;
    const int ns = (int)sidx_.size();
    #pragma omp parallel for schedule(dynamic, 64) reduction(+ : nnz)
    for (int i = 0; i < npos_; i++) {
        size_t idx1 = cidx_[i];
        double r1 = rdi_[i];
        double x1 = pos0_[3 * i], y1 = pos0_[3 * i + 1], z1 = pos0_[3 * i + 2];
        int ix1 = (int)(idx1 / nc2), iy1 = (int)((idx1 % nc2) / nc1), iz1 = (int)(idx1 % nc1);
        std::vector<int>& row = pairs_[i];
        row.clear();
        for (int k = 0; k < ns; k++) {
            int ix2 = wrap_cell(ix1 + sidx_[k].dx, nc1);
            int iy2 = wrap_cell(iy1 + sidx_[k].dy, nc1);
            int iz2 = wrap_cell(iz1 + sidx_[k].dz, nc1);
            size_t idx2 = ix2 * nc2 + iy2 * nc1 + iz2;
            int j = head_[idx2];
            double max_sum = sidx_[k].max_sum, min_sum = sidx_[k].min_sum;
            if (r1 + min_radius_ > max_sum) {
                for (; j != -1; j = next_[j]) row.push_back(j);
            } else if (r1 + max_radius_ > min_sum) {
                for (; j != -1; j = next_[j]) {
                    double r2 = rdi_[j];
                    if (r1 + r2 > max_sum) {
                        row.push_back(j);
                    } else if (r1 + r2 > min_sum) {
                        double rx = fabs(pos0_[3 * j] - x1);
                        double ry = fabs(pos0_[3 * j + 1] - y1);
                        double rz = fabs(pos0_[3 * j + 2] - z1);
                        rx = (rx > box_ / 2.0 ? rx - box_ : rx);
                        ry = (ry > box_ / 2.0 ? ry - box_ : ry);
                        rz = (rz > box_ / 2.0 ? rz - box_ : rz);
                        double rr = rx * rx + ry * ry + rz * rz;
                        if (rr == 0.0) rr = 1.0e-10;
                        double s = 2.0 * sqrt(rr) / (r1 + r2);
                        if (s < cutoff_) row.push_back(j);
                    }
                }
            }
        }
        std::sort(row.begin(), row.end());
        nnz += (int)row.size();
    }
    return nnz;
}

// src/pairlist/pair_base.cc:118-123
int PairList::build(const double* pos) {
    construct_cells(pos);
    return mode_ == 'R' ? find_pairs_r() : find_pairs_s();
}

// src/pairlist/pair_base.cc:126-140
void PairList::get_pairs(int* rowptr, int* colidx) const {
    rowptr[0] = 0;
    for (int i = 0; i < npos_; i++) rowptr[i + 1] = rowptr[i] + (int)pairs_[i].size();
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < npos_; i++) {
        int s = rowptr[i];
        for (size_t j = 0; j < pairs_[i].size(); j++) colidx[s + j] = pairs_[i][j];
    }
}

}  // namespace orc
