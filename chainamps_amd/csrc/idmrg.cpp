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
//           SquareRescale(1), MakePrediction (DMRG_routines.cpp:693-707);
//   finite sweeps: SuperBlock::move_right_two_vertex / move_left_two_vertex
//           (DMRG_routines.cpp:376-444) with MakeRFinitePrediction /
//           MakeLFinitePrediction (709-716) and the finite CheckConvergence
//           (773-776), driven by FiniteDMRG::run (fDMRG_routines.cpp:19-66).
// Every tensor stays on the device: the reference's disk block store
// (BlocksStructure, DMRG_routines.cpp:57-186: one .BLOCK file per block
// size, one .MPS_matrix file per vertex) is a set of device-resident
// tensors indexed the same way.
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

TensorPtr make_l_finite_prediction(Ctx& ctx, const BSTensor& U, const std::vector<double>& values,
                                   const BSTensor& A) {
  // A . (U . diag(values)): A's right bond against U's left bond
  TensorPtr Us = copy_tensor(ctx, U);
  scale_along_index(ctx, *Us, 2, values);
  return contract(ctx, A, false, *Us, false, {{2, 1}});
}

TensorPtr make_r_finite_prediction(Ctx& ctx, const std::vector<double>& values, const BSTensor& Vd,
                                   const BSTensor& B) {
  // reorder(diag(values) . Vd, 102, 2) . B: Vd's right bond against B's left bond
  TensorPtr Vs = copy_tensor(ctx, Vd);
  scale_along_index(ctx, *Vs, 0, values);
  TensorPtr T = reorder(ctx, *Vs, false, {1, 0, 2}, 2);
  return contract(ctx, *T, false, B, false, {{2, 0}});
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
  CAM_REQUIRE(ncv >= 2 && ncv <= 63, "iDMRG: ncv must be 2..63");
}

namespace {
template <typename T>
void set_at(std::vector<T>& v, int64_t k, T x) {
  if (static_cast<int64_t>(v.size()) <= k) v.resize(k + 1);
  v[k] = std::move(x);
}
}  // namespace

IDMRGRecord IDMRG::grow(int64_t chi, double trunc) {
  CAM_REQUIRE(!swept, "iDMRG: growth after finite sweeps is not supported (FiniteDMRG does not grow)");
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
    lblk = {LBd, LB};
    rblk = {RBd, RB};
    lmps = {nullptr, U};
    rmps = {nullptr, Vd};
    nleft = nright = 0;
  } else {
    if (size > 2) {
      LB = make_left_block(*ctx, *LB, *W, *U);
      RB = make_right_block(*ctx, *RB, *W, *Vd);
      set_at(lblk, size / 2, LB);
      set_at(rblk, size / 2, RB);
    }
    size += 2;
    nleft = nright = size / 2 - 1;
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
    set_at(lmps, size / 2, U);
    set_at(rmps, size / 2, Vd);
  }
  square_rescale(values);
  // MakePrediction maps Vd's right bond onto U's left bond; when the two
  // carry different charges (a non-zero target) there is no prediction and
  // the next solve starts from a random vector, which the record reports.
  // Every other failure propagates.
  if (Vd->idx[2].same_charges(U->idx[1])) {
    guess = make_prediction(*ctx, *U, values, *Vd, prev_lambda);
    guess_dropped = false;
  } else {
    guess.reset();
    guess_dropped = true;
  }
  // 1 - fidelity of the prediction (CheckConvergence, DMRG_routines.cpp:769-771)
  rec.fidelity = fidelity(*ctx, values, *Vd, prev_lambda);
  rec.guess_dropped = guess_dropped ? 1.0 : 0.0;
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
  rec.left = nleft;
  rec.right = nright;
  return rec;
}

// SquareRescale(1.0) (MPS_matrix.hpp:66-70): values to unit square sum
void IDMRG::square_rescale(std::vector<double>& v) {
  double sq = 0.0;
  for (double x : v) sq += x * x;
  CAM_REQUIRE(sq > 0.0, "Square sum error");
  for (double& x : v) x *= std::sqrt(1.0 / sq);
}

IDMRGRecord IDMRG::move(bool to_right, int64_t chi, double trunc) {
  CAM_REQUIRE(size >= 4, "finite sweeps need at least four vertices (two are stored as they are)");
  const auto t0 = std::chrono::steady_clock::now();
  auto lap = [this](std::chrono::steady_clock::time_point& t) {
    HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const auto now = std::chrono::steady_clock::now();
    const double dt = std::chrono::duration<double>(now - t).count();
    t = now;
    return dt;
  };
  auto tp = t0;
  IDMRGRecord rec{};
  // shift the centre: the vertex that leaves it joins the block on the side
  // it leaves towards, the stored MPS matrix of the vertex that enters it
  // gives the prediction (the truncation is switched off next to the ends)
  TensorPtr pred;
  double t = trunc;
  if (to_right) {
    CAM_REQUIRE(nright >= 1, "At right end of superblock already...");
    TensorPtr nl = make_left_block(*ctx, *lblk[nleft], *W, *U);
    ++nleft;
    --nright;
    set_at(lblk, nleft, nl);
    CAM_REQUIRE(static_cast<int64_t>(rmps.size()) > nright + 1 && rmps[nright + 1], "finite DMRG: no stored right matrix");
    pred = make_r_finite_prediction(*ctx, values, *Vd, *rmps[nright + 1]);
    if (nright <= 1) t = -0.0;
  } else {
    CAM_REQUIRE(nleft >= 1, "At left end of superblock already...");
    TensorPtr nr = make_right_block(*ctx, *rblk[nright], *W, *Vd);
    ++nright;
    --nleft;
    set_at(rblk, nright, nr);
    CAM_REQUIRE(static_cast<int64_t>(lmps.size()) > nleft + 1 && lmps[nleft + 1], "finite DMRG: no stored left matrix");
    pred = make_l_finite_prediction(*ctx, *U, values, *lmps[nleft + 1]);
    if (nleft <= 1) t = -0.0;
  }
  swept = true;
  LB = lblk[nleft];
  RB = rblk[nright];
  prev_lambda = values;
  Heff H;
  H.build(*ctx, LB, W, RB);
  xbuf.ensure(std::max<int64_t>(H.n, 1));
  gbuf.ensure(std::max<int64_t>(H.n, 1));
  H.vector_from_tensor(*pred, gbuf.get());
  launch_copy(gbuf.get(), xbuf.get(), H.n, ctx->stream);
  rec.block_seconds += lap(tp);
  double info[4] = {0, 0, 0, 0};
  const double E = eigs_lowest(H, xbuf.get(), true, converge, 0.0, ncv, mixed, info);
  rec.eig_seconds = lap(tp);
  TensorPtr psi = H.tensor_from_vector(xbuf.get());
  double disc = 0.0;
  svd_truncate(*ctx, *psi, chi, t, U, Vd, values, disc);
  truncation = disc;
  square_rescale(values);
  set_at(lmps, nleft + 1, U);
  set_at(rmps, nright + 1, Vd);
  rec.svd_seconds = lap(tp);
  // CheckConvergence(Prediction, MPSDecomposition): 1 - |<guess|U s Vd>|
  TensorPtr Us = copy_tensor(*ctx, *U);
  scale_along_index(*ctx, *Us, 2, values);
  TensorPtr now = contract(*ctx, *Us, false, *Vd, false, {{2, 0}});
  H.vector_from_tensor(*now, xbuf.get());
  DevBuf<cplx> part(static_cast<size_t>(multi_dot_blocks(H.n))), c(1);
  launch_multi_dot(gbuf.get(), H.n, 1, xbuf.get(), H.n, part.get(), c.get(), ctx->stream);
  cplx ov{0, 0};
  HIP_CHECK(hipMemcpyAsync(&ov, c.get(), sizeof(cplx), hipMemcpyDeviceToHost, ctx->stream));
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  rec.fidelity = 1.0 - std::hypot(ov.re, ov.im);
  guess.reset();
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
  rec.left = nleft;
  rec.right = nright;
  return rec;
}

std::vector<IDMRGRecord> IDMRG::finite_sweeps(int64_t num_sweeps, int64_t chi, double trunc,
                                               std::vector<int>* midpoint) {
  std::vector<IDMRGRecord> out;
  if (size <= 2) return out;   // two vertices: nothing to sweep
  auto push = [&](const IDMRGRecord& r, bool record_mid) {
    out.push_back(r);
    if (midpoint) midpoint->push_back(record_mid && nleft == nright ? 1 : 0);
  };
  for (int64_t sweep = 0; sweep < num_sweeps; ++sweep) {
    for (int64_t r = nright; r > 0; --r) push(move(true, chi, trunc), false);
    for (int64_t l = nleft; l > 0; --l) push(move(false, chi, trunc), true);
    // for (r = right_size(); r > left_size(); --r): both sides move
    for (int64_t r = nright; r > nleft; --r) push(move(true, chi, trunc), true);
  }
  return out;
}

}  // namespace cam
