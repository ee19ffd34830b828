// Oracle periodic cell-list pair lists (test infrastructure only; see
// orc_common.hpp). Restates PairListBase/PairListR/PairListS
// (src/pairlist/pair_base.cc:34-140, pair_r.cc:15-170, pair_s.cc:16-186).
#pragma once
#include <vector>

namespace orc {

class PairList {
  public:
    // mode 'R': pairs with min-image r <= cutoff (self included);
    // mode 'S': pairs with 2r/(a_i+a_j) < cutoff (self included).
    PairList(int npos, const double* rdi, double box, double cutoff, char mode);
    int init();                      // estimated nnz, 0 on invalid input
    int build(const double* pos);    // exact nnz
    void get_pairs(int* rowptr, int* colidx) const;
    int nc1() const { return nc1_; }

  private:
    int num_cells() const;
    void init_search_idx();
    int init_pairs();
    void construct_cells(const double* pos);
    int find_pairs_r();
    int find_pairs_s();

    struct Sidx { int dx, dy, dz; bool all; double min_sum, max_sum; };
    int npos_;
    double box_, cutoff_;
    char mode_;
    std::vector<double> rdi_;
    int nc1_ = 0;
    double cell_size_ = 0;
    double max_radius_ = 0, min_radius_ = 0;
    std::vector<Sidx> sidx_;
    std::vector<int> head_, next_;
    std::vector<size_t> cidx_;
    std::vector<double> pos0_;
    std::vector<std::vector<int>> pairs_;
};

}  // namespace orc
