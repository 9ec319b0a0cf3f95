// Device-resident infinite DMRG growth (SuperBlock / iDMRG).
//
// Mirrors SuperBlock::initialise and SuperBlock::grow_two_vertex
// (DMRG_routines.cpp:309-374) and iDMRG::run (iDMRG_routines.cpp:19-33):
//   step 0: the two-vertex problem with the dummy end blocks
//           (MakeDummyLeftBlock/RightBlock, DMRG_routines.cpp:613-653 -- the
//           open-boundary MPOs LeftH/RightH are W with those ends
//           contracted), solved to convergence (the reference uses a dense
//           Hermitian solve here), TwoVertexSVD with bond dimension d when
//           chi < 1 and no truncation is requested, SquareRescale(1), the
//           first left/right blocks;
//   step k: MakeLeftBlock/MakeRightBlock from the previous U / V^dagger
//           (skipped at size 2, whose blocks were made at step 0),
//           TwoVertexWavefunction with the predicted guess, TwoVertexSVD,
//           SquareRescale(1), MakePrediction (DMRG_routines.cpp:693-707).
// Every tensor stays on the device; the disk block store of the reference
// (BlocksStructure, DMRG_routines.cpp:57-186) is not reproduced.
#include <chrono>
#include <cmath>

#include "idmrg.hpp"

namespace cam {

namespace {

TensorPtr small_tensor(Ctx& ctx, const Rules& rules, std::vector<QIndex> idx, int numrows,
                       int64_t row, int64_t col) {
  int64_t rows = 1, cols = 1;
  for (int k = 0; k < numrows; ++k) rows *= idx[k].size;
  for (size_t k = numrows; k < idx.size(); ++k) cols *= idx[k].size;
  std::vector<int64_t> p(cols + 1, 0);
  for (int64_t c = col + 1; c <= cols; ++c) p[c] = 1;
  const int64_t i = row;
  const double x[2] = {1.0, 0.0};
  return tensor_from_csc(rules, std::move(idx), numrows, rows, cols, p.data(), &i, x, ctx.stream);
}

QIndex scalar_index(bool inward, const Charge& q) {
  QIndex r;
  r.inward = inward;
  r.size = 1;
  r.charges = {q};
  return r;
}

}  // namespace

double entropy(const std::vector<double>& v) {
  double s = 0.0;
  for (double x : v)
    if (x != 0.0) s -= x * x * std::log(x * x);
  return s;
}

TensorPtr make_dummy_left_block(Ctx& ctx, const BSTensor& W) {
  const Charge z = W.rules.zero();
  std::vector<QIndex> idx{scalar_index(false, z), scalar_index(true, z), scalar_index(true, z),
                          W.idx[1].flipped(), scalar_index(true, z), scalar_index(false, z)};
  idx[3].inward = false;
  idx[3].physical = false;
  return small_tensor(ctx, W.rules, idx, 4, W.idx[1].size - 1, 0);
}

TensorPtr make_dummy_right_block(Ctx& ctx, const BSTensor& W, const Charge& target) {
  const Charge z = W.rules.zero();
  QIndex b = W.idx[3];
  b.inward = true;
  b.physical = false;
  std::vector<QIndex> idx{scalar_index(false, target), scalar_index(true, z), b,
                          scalar_index(false, z), scalar_index(true, target), scalar_index(false, z)};
  return small_tensor(ctx, W.rules, idx, 3, 0, 0);
}

TensorPtr make_left_block(Ctx& ctx, const BSTensor& LB, const BSTensor& W, const BSTensor& A) {
  TensorPtr t = contract(ctx, LB, false, A, false, {{5, 1}});
  t = contract(ctx, W, false, *t, false, {{1, 3}, {2, 5}});
  t = contract(ctx, A, true, *t, false, {{0, 0}, {1, 3}});
  t->move_dummy_indices({2, 0, 3, 1, 4, 5});
  return t;
}

TensorPtr make_right_block(Ctx& ctx, const BSTensor& RB, const BSTensor& W, const BSTensor& B) {
  TensorPtr t = contract(ctx, RB, false, B, false, {{4, 2}});
  t = contract(ctx, W, false, *t, false, {{2, 6}, {3, 2}});
  t = contract(ctx, B, true, *t, false, {{1, 0}, {2, 2}});
  t->move_dummy_indices({0, 2, 1, 3, 5, 4});
  return t;
}

static TensorPtr copy_tensor(Ctx& ctx, const BSTensor& A) {
  auto t = std::make_shared<BSTensor>();
  t->rules = A.rules;
  t->idx = A.idx;
  t->numrows = A.numrows;
  t->blocks = A.blocks;
  t->finalize_layout();
  if (A.total)
    HIP_CHECK(hipMemcpyAsync(t->data.get(), A.data.get(), A.total * sizeof(cplx), hipMemcpyDeviceToDevice,
                             ctx.stream));
  return t;
}

TensorPtr make_prediction(Ctx& ctx, const BSTensor& U, const std::vector<double>& values, const BSTensor& Vd,
                          const std::vector<double>& prev) {
  // LambdaLB = reorder(U . diag(values) on the new bond, 102, 1)
  TensorPtr Us = copy_tensor(ctx, U);
  scale_along_index(ctx, *Us, 2, values);
  TensorPtr LambdaLB = reorder(ctx, *Us, false, {1, 0, 2}, 1);
  // ALambdaR = reorder(diag(values) . Vd, 102, 2)
  TensorPtr Vs = copy_tensor(ctx, Vd);
  scale_along_index(ctx, *Vs, 0, values);
  TensorPtr ALambdaR = reorder(ctx, *Vs, false, {1, 0, 2}, 2);
  // . diag(1/prev) from Vd's right bond to U's left bond (same positions)
  CAM_REQUIRE(ALambdaR->idx[2].same_charges(U.idx[1]),
              "prediction: previous left and right bonds carry different charges");
  std::vector<double> inv(prev.size());
  for (size_t i = 0; i < prev.size(); ++i) inv[i] = 1.0 / prev[i];
  scale_along_index(ctx, *ALambdaR, 2, inv);
  ALambdaR->idx[2] = U.idx[1];
  ALambdaR->idx[2].inward = false;
  ALambdaR->rebuild_lookup();
  return contract(ctx, *ALambdaR, false, *LambdaLB, false, {{2, 0}});
}

double fidelity(Ctx& ctx, const std::vector<double>& values, const BSTensor& Vd, const std::vector<double>& prev) {
  // MakePrediction's Auxiliary = diag(s') W^dagger comes from the SVD
  // ALambdaR = reorder(diag(values) . Vd, 102, 2) = C s' W^dagger, so
  // Auxiliary . diag(prev) = C^dagger (ALambdaR . diag(prev)) and, C being an
  // isometry, both have the same singular values: CheckConvergence is
  // 1 - sum of the singular values of ALambdaR . diag(prev), values only.
  TensorPtr Vs = copy_tensor(ctx, Vd);
  scale_along_index(ctx, *Vs, 0, values);
  CAM_REQUIRE(static_cast<int64_t>(prev.size()) == Vs->idx[2].size, "fidelity: previous bond size mismatch");
  scale_along_index(ctx, *Vs, 2, prev);
  TensorPtr X = reorder(ctx, *Vs, false, {1, 0, 2}, 2);
  std::vector<double> sv;
  singular_values(ctx, *X, sv);
  double sum = 0.0;
  for (double x : sv) sum += x;
  return 1.0 - sum;
}

IDMRG::IDMRG(Ctx& c, TensorPtr mpo, const Charge& tgt, bool conv, bool mix, int k)
    : ctx(&c), W(std::move(mpo)), target(tgt), converge(conv), mixed(mix), ncv(k) {
  CAM_REQUIRE(W->nidx() == 4, "iDMRG needs a 4-index MPO");
}

IDMRGRecord IDMRG::grow(int64_t chi, double trunc) {
  const auto t0 = std::chrono::steady_clock::now();
  // phase clock: the stream is drained so each phase is charged its own work
  auto lap = [this](std::chrono::steady_clock::time_point& t) {
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const auto now = std::chrono::steady_clock::now();
    const double dt = std::chrono::duration<double>(now - t).count();
    t = now;
    return dt;
  };
  auto tp = t0;
  IDMRGRecord rec{};
  Heff H;
  double info[4] = {0, 0, 0, 0};
  double E = 0.0;
  if (size == 0) {
    LBd = make_dummy_left_block(*ctx, *W);
    RBd = make_dummy_right_block(*ctx, *W, target);
    H.build(*ctx, LBd, W, RBd);
    xbuf.ensure(std::max<int64_t>(H.n, 1));
    rec.block_seconds += lap(tp);
    E = eigs_lowest(H, xbuf.get(), false, true, 0.0, ncv, false, info);
    rec.eig_seconds = lap(tp);
    TensorPtr psi = H.tensor_from_vector(xbuf.get());
    const int64_t bond = (chi < 1 && trunc <= 0.0) ? W->idx[0].size : chi;
    double disc = 0.0;
    svd_truncate(*ctx, *psi, bond, trunc, U, Vd, values, disc);
    truncation = disc;
    rec.svd_seconds = lap(tp);
    LB = make_left_block(*ctx, *LBd, *W, *U);
    RB = make_right_block(*ctx, *RBd, *W, *Vd);
    size = 2;
    prev_lambda = {1.0};
  } else {
    if (size > 2) {
      LB = make_left_block(*ctx, *LB, *W, *U);
      RB = make_right_block(*ctx, *RB, *W, *Vd);
    }
    size += 2;
    prev_lambda = values;
    H.build(*ctx, LB, W, RB);
    xbuf.ensure(std::max<int64_t>(H.n, 1));
    bool have = false;
    if (guess) {
      H.vector_from_tensor(*guess, xbuf.get());
      have = true;
    }
    rec.block_seconds += lap(tp);
    E = eigs_lowest(H, xbuf.get(), have, converge, 0.0, ncv, mixed, info);
    rec.eig_seconds = lap(tp);
    TensorPtr psi = H.tensor_from_vector(xbuf.get());
    double disc = 0.0;
    svd_truncate(*ctx, *psi, chi, trunc, U, Vd, values, disc);
    truncation = disc;
    rec.svd_seconds = lap(tp);
  }
  // SquareRescale(1.0) (MPS_matrix.hpp:66-70): Truncation/1, values to unit square sum
  double sq = 0.0;
  for (double v : values) sq += v * v;
  CAM_REQUIRE(sq > 0.0, "Square sum error");
  for (double& v : values) v *= std::sqrt(1.0 / sq);
  try {
    guess = make_prediction(*ctx, *U, values, *Vd, prev_lambda);
  } catch (const Error&) {
    guess.reset();   // inconsistent bond charges: next solve starts at random
  }
  // 1 - fidelity of the prediction (CheckConvergence, DMRG_routines.cpp:769-771)
  rec.fidelity = fidelity(*ctx, values, *Vd, prev_lambda);
  rec.block_seconds += lap(tp);
  rec.N = size;
  rec.energy = E;
  rec.energy_per_vertex = E / static_cast<double>(size);
  rec.entropy = entropy(values);
  rec.truncation = truncation;
  rec.chi = static_cast<int64_t>(values.size());
  rec.n = H.n;
  rec.matvecs_fp64 = info[0];
  rec.matvecs_fp32 = info[1];
  rec.residual = info[3];
  rec.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  rec.heff_flops = H.flops;
  return rec;
}

}  // namespace cam
