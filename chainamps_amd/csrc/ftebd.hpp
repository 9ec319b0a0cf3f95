// Finite-chain TEBD with the MPS resident in device memory (the TEBD class
// of TEBD_routines.cpp:273-791, whose site matrices live on disk).
#pragma once
#include <vector>

#include "solvers.hpp"

namespace cam {

// TrotterDecomposition (TEBD_routines.cpp:14-85): the bond operators of one
// time step and the order in which they are applied (odd bonds at even
// positions of `seq`), for orders 1, 2 and 4.
void trotter_operators(Ctx& ctx, const BSTensor& W, double tau, int order, std::vector<TensorPtr>& ops,
                       std::vector<int>& seq);

struct FTEBDRecord {
  int64_t step;
  double time, truncation, entropy, overlap_re, overlap_im, abs_overlap;
};

struct FTEBD {
  Ctx* ctx;
  int order;
  double dt;
  int64_t N;
  int64_t step = 0;
  std::vector<TensorPtr> ops;
  std::vector<int> seq;
  // site v = 1..N (entry 0 unused): left shape [s, a_l, a_r] (two row
  // indices), right shape [a_l, s, a_r] (one), initial state for overlaps
  std::vector<TensorPtr> left, right, initial;
  double max_truncation = 0.0;
  std::vector<FTEBDRecord> records;

  // `sites`: a left-canonical MPS (FiniteMPS::makeLC) with an even number
  // of vertices >= 2
  FTEBD(Ctx& c, const BSTensor& W, const std::vector<TensorPtr>& sites, double timestep, int trotter_order);
  void apply_to_odd_bonds(const BSTensor& G, int64_t D, double minS);
  void apply_to_even_bonds(const BSTensor& G, int64_t D, double minS);
  void left_canonise(bool measure);
  void evolve(int64_t steps, int64_t D, double minS, int64_t interval);
};

}  // namespace cam
