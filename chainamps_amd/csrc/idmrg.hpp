// Device-resident iDMRG growth (SuperBlock::grow_two_vertex mirror).
#pragma once
#include <vector>

#include "solvers.hpp"

namespace cam {

struct IDMRGRecord {
  int64_t N;
  double energy, energy_per_vertex, entropy, truncation;
  int64_t chi, n;
  double matvecs_fp64, matvecs_fp32, residual, seconds, heff_flops;
  double eig_seconds, svd_seconds, block_seconds;   // phases of `seconds`
  double fidelity;                                  // 1 - fidelity of the prediction
};

TensorPtr make_dummy_left_block(Ctx& ctx, const BSTensor& W);
TensorPtr make_dummy_right_block(Ctx& ctx, const BSTensor& W, const Charge& target);
TensorPtr make_left_block(Ctx& ctx, const BSTensor& LB, const BSTensor& W, const BSTensor& A);
TensorPtr make_right_block(Ctx& ctx, const BSTensor& RB, const BSTensor& W, const BSTensor& B);
// CheckConvergence (DMRG_routines.cpp:769-771) of the prediction built from
// (values, Vd) against the previous step's bond weights: 1 - sum of the
// singular values of Auxiliary . diag(prev)
double fidelity(Ctx& ctx, const std::vector<double>& values, const BSTensor& Vd, const std::vector<double>& prev);
TensorPtr make_prediction(Ctx& ctx, const BSTensor& U, const std::vector<double>& values, const BSTensor& Vd,
                          const std::vector<double>& prev);

struct IDMRG {
  Ctx* ctx;
  TensorPtr W;
  Charge target;
  bool converge, mixed;
  int ncv;
  int64_t size = 0;
  TensorPtr LBd, RBd, LB, RB, U, Vd, guess;
  std::vector<double> values, prev_lambda;
  double truncation = 0.0;
  DevBuf<cplx> xbuf;   // eigenvector of the current step (grow-only)

  IDMRG(Ctx& c, TensorPtr mpo, const Charge& tgt, bool conv, bool mix, int k);
  IDMRGRecord grow(int64_t chi, double trunc);
};

}  // namespace cam
