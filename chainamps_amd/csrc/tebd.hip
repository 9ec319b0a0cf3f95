// iTEBD bond update on the device: iTEBD::apply_and_decompose
// (TEBD_routines.cpp:233-266).
//
//   C   = reorder(contract(G, contract(A0, A1, 21), 2032), 0213, rows 2)
//   SVD of C * diag(lambda0) on its last index, rows (s1', a_l)
//   A0' = U;  A1' = reorder(contract(U^dagger, C, 0011), 102, rows 2)
//                   / sqrt(sum lambda0^2)
//   unit-cell swap: the returned pair is (A1', A0'), the new weights are
//   the singular values, the other bond keeps lambda0.
// G is the 4-index bond operator [s1'(in), s2'(in), s1(out), s2(out)]
// (MakeBondEvolutionOperator, TEBD_routines.cpp:869-880); A0, A1 are
// left-shape MPS tensors [s(in), a_l(in), a_r(out)].
#include <cmath>

#include "solvers.hpp"

namespace cam {

namespace {

struct ScaleJob {
  double2* p;
  int64_t size, stride, dim;
  const double* lam;       // per-sector weights (dim entries)
};

__global__ void scale_index_kernel(const ScaleJob* jobs, int njobs) {
  for (int j = blockIdx.y; j < njobs; j += gridDim.y) {
    const ScaleJob J = jobs[j];
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < J.size;
         e += int64_t(gridDim.x) * blockDim.x) {
      const double f = J.lam[(e / J.stride) % J.dim];
      double2 v = J.p[e];
      v.x *= f;
      v.y *= f;
      J.p[e] = v;
    }
  }
}

}  // namespace

// Multiply a tensor along index k by the diagonal `lam` (one value per
// position of that index): contract(T, diag MPX, (k,0)) for k the last
// index keeps the index order (MPX_matrix.cpp:44-54 diagonal constructor).
void scale_along_index(Ctx& ctx, BSTensor& T, int k, const std::vector<double>& lam) {
  CAM_REQUIRE(static_cast<int64_t>(lam.size()) == T.idx[k].size, "diagonal size mismatch");
  const QIndex& q = T.idx[k];
  std::vector<double> per;
  std::vector<int64_t> soff(q.nsec());
  for (int s = 0; s < q.nsec(); ++s) {
    soff[s] = static_cast<int64_t>(per.size());
    for (auto p : q.sec_pos[s]) per.push_back(lam[p]);
  }
  DevBuf<double> dl(std::max<size_t>(per.size(), 1));
  dl.upload(per.data(), per.size(), ctx.stream);
  std::vector<ScaleJob> jobs;
  for (const Block& b : T.blocks) {
    int64_t st = 1;
    for (int i = 0; i < k; ++i) st *= b.dims[i];
    if (b.size)
      jobs.push_back(ScaleJob{reinterpret_cast<double2*>(T.data.get() + b.offset), b.size, st, b.dims[k],
                              dl.get() + soff[b.sec[k]]});
  }
  if (jobs.empty()) return;
  DevBuf<ScaleJob> dj(jobs.size());
  dj.upload(jobs.data(), jobs.size(), ctx.stream);
  hipLaunchKernelGGL(scale_index_kernel, dim3(64, std::min<size_t>(jobs.size(), 1024)), dim3(256), 0,
                     ctx.stream, dj.get(), static_cast<int>(jobs.size()));
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(ctx.stream));
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

void tebd_gate(Ctx& ctx, const BSTensor& G, const BSTensor& A0, const BSTensor& A1,
               const std::vector<double>& lambda0, int64_t D, double minS, TensorPtr& A0n,
               TensorPtr& A1n, std::vector<double>& lambda1, double& truncation) {
  CAM_REQUIRE(G.nidx() == 4 && A0.nidx() == 3 && A1.nidx() == 3, "tebd_gate: bad tensor ranks");
  TensorPtr X = contract(ctx, A0, false, A1, false, {{2, 1}});
  TensorPtr Y = contract(ctx, G, false, *X, false, {{2, 0}, {3, 2}});
  TensorPtr C = reorder(ctx, *Y, false, {0, 2, 1, 3}, 2);
  TensorPtr CL = copy_tensor(ctx, *C);
  scale_along_index(ctx, *CL, 3, lambda0);
  CL->numrows = 2;                                   // ShiftNumRowIndices(2)
  TensorPtr U, Vd;
  std::vector<double> vals;
  double disc = 0.0;
  svd_truncate(ctx, *CL, D, minS, U, Vd, vals, disc);
  TensorPtr A1p = contract(ctx, *U, true, *C, false, {{0, 0}, {1, 1}});
  A1p = reorder(ctx, *A1p, false, {1, 0, 2}, 2);
  double sq = 0.0;
  for (double l : lambda0) sq += l * l;
  CAM_REQUIRE(sq > 0.0, "tebd_gate: zero lambda0");
  launch_scale(cplx{1.0 / std::sqrt(sq), 0.0}, A1p->data.get(), A1p->total, ctx.stream);
  HIP_CHECK(hipStreamSynchronize(ctx.stream));
  // unit-cell swap(0,1)
  A0n = A1p;
  A1n = U;
  lambda1 = vals;
  truncation = disc;
}

}  // namespace cam
