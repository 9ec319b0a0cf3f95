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
  double guess_dropped;                             // 1: no prediction for the next step
  int64_t left, right;                              // vertices in the left / right block
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
// MakeLFinitePrediction (DMRG_routines.cpp:709-712): A . (U diag(values)),
// the wavefunction [s1, a0, s2, a2] of a step to the left
TensorPtr make_l_finite_prediction(Ctx& ctx, const BSTensor& U, const std::vector<double>& values,
                                   const BSTensor& A);
// MakeRFinitePrediction (DMRG_routines.cpp:713-716): reorder(diag(values)
// Vd, 102, 2) . B, the wavefunction of a step to the right
TensorPtr make_r_finite_prediction(Ctx& ctx, const std::vector<double>& values, const BSTensor& Vd,
                                   const BSTensor& B);

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
  bool guess_dropped = false;   // the last step made no prediction
  DevBuf<cplx> xbuf;   // eigenvector of the current step (grow-only)
  DevBuf<cplx> gbuf;   // prediction of a finite step (grow-only)
  // Block store (BlocksStructure, DMRG_routines.cpp:57-186), device
  // resident: lblk[k] / rblk[k] hold k vertices (k = 0: the dummy end
  // blocks), lmps[k] / rmps[k] are the MPS matrices of vertex k counted
  // from the left / right end; nleft / nright are the block sizes of the
  // current centre.
  std::vector<TensorPtr> lblk, rblk, lmps, rmps;
  int64_t nleft = 0, nright = 0;
  bool swept = false;

  IDMRG(Ctx& c, TensorPtr mpo, const Charge& tgt, bool conv, bool mix, int k);
  IDMRGRecord grow(int64_t chi, double trunc);
  // one finite step: SuperBlock::move_right_two_vertex (to_right) or
  // move_left_two_vertex (DMRG_routines.cpp:376-444)
  IDMRGRecord move(bool to_right, int64_t chi, double trunc);
  // FiniteDMRG::run (fDMRG_routines.cpp:19-66); midpoint (if given) flags
  // the records the reference pushes to its output (centre in the middle
  // on the way left and on the way back)
  std::vector<IDMRGRecord> finite_sweeps(int64_t num_sweeps, int64_t chi, double trunc,
                                         std::vector<int>* midpoint = nullptr);

 private:
  static void square_rescale(std::vector<double>& v);
};

}  // namespace cam
