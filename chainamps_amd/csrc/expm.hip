// Gate construction for (i)TEBD on the device.
//
//   * ExtractMPOBlock (MPO_matrix.cpp:69-85): the sub-MPO of rows
//     [r0, r1] of the left MPO bond and columns [c0, c1] of the right one;
//   * MakeOddBondHamiltonian / MakeEvenBondHamiltonian
//     (TEBD_routines.cpp:826-867): the last row of W times its first
//     column (odd: all of the inner bond; even: the coupling channels
//     1..w-2 only), contracted over the inner bond (3,1) and reordered
//     032415 with two row indices: [s1', s2', s1, s2, dummy, dummy];
//   * MakeBondEvolutionOperator (TEBD_routines.cpp:869-880) with
//     Exponentiate (sparse_interface.cpp:1707-1714) over the blockwise
//     Hermitian decomposition SparseMatrix::HED (sparse_interface.cpp:
//     1414-1441): per group of columns with equal outgoing charge, the
//     full square block (zero eigenvalues kept) is diagonalised
//     (hermitian_eig), eigenvector entries |v| <= eps/rows are dropped, and
//     G = V diag(exp(-i tau lambda)) V^dagger (rocBLAS zgemm).  The result
//     keeps the first four indices of the bond Hamiltonian.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>

#include "solvers.hpp"

namespace cam {

namespace {

__global__ void map_gather_kernel(const double2* __restrict__ src, const int64_t* __restrict__ map,
                                  double2* __restrict__ dst, int64_t n) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t m = map[e];
    dst[e] = m >= 0 ? src[m] : make_double2(0.0, 0.0);
  }
}

// V[:, j] entries with |v| <= thr set to zero; B[:, j] = V[:, j] exp(f lambda_j)
__global__ void exp_columns_kernel(double2* __restrict__ V, double2* __restrict__ B, const double* __restrict__ lam,
                                   int64_t n, double fre, double fim, double thr) {
  const int64_t tot = n * n;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < tot;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t j = e / n;
    double2 v = V[e];
    if (sqrt(v.x * v.x + v.y * v.y) <= thr) v = make_double2(0.0, 0.0);
    V[e] = v;
    const double a = fre * lam[j], b = fim * lam[j];
    const double ea = exp(a);
    const double er = ea * cos(b), ei = ea * sin(b);
    B[e] = make_double2(v.x * er - v.y * ei, v.x * ei + v.y * er);
  }
}

int grid_of(int64_t n) { return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256))); }

// global position of every element of block b, per index
void block_positions(const BSTensor& T, const Block& b, int64_t e, int64_t* g) {
  for (int k = 0; k < T.nidx(); ++k) {
    const int64_t l = e % b.dims[k];
    e /= b.dims[k];
    g[k] = T.idx[k].sec_pos[b.sec[k]][l];
  }
}

}  // namespace

TensorPtr extract_mpo_block(Ctx& ctx, const BSTensor& W, int64_t r0, int64_t r1, int64_t c0, int64_t c1) {
  CAM_REQUIRE(W.nidx() == 4 && W.numrows == 2, "ExtractMPOBlock needs a 4-index MPO with two row indices");
  const int64_t d0 = W.idx[0].size, wl = W.idx[1].size, d2 = W.idx[2].size, wr = W.idx[3].size;
  CAM_REQUIRE(0 <= r0 && r0 <= r1 && r1 < wl && 0 <= c0 && c0 <= c1 && c1 < wr, "ExtractMPOBlock: range out of bounds");
  hipStream_t s = ctx.stream;
  const int64_t nnz = tensor_to_csc(W, 0, nullptr, nullptr, nullptr, s);
  std::vector<int64_t> p(W.cols() + 1), i(std::max<int64_t>(nnz, 1));
  std::vector<double> x(2 * std::max<int64_t>(nnz, 1));
  tensor_to_csc(W, nnz, p.data(), i.data(), x.data(), s);
  const int64_t nr = r1 - r0 + 1, nc = c1 - c0 + 1;
  std::vector<QIndex> idx = W.idx;
  idx[1].size = nr;
  idx[1].charges.assign(W.idx[1].charges.begin() + r0, W.idx[1].charges.begin() + r1 + 1);
  idx[1].build_sectors();
  idx[3].size = nc;
  idx[3].charges.assign(W.idx[3].charges.begin() + c0, W.idx[3].charges.begin() + c1 + 1);
  idx[3].build_sectors();
  const int64_t rows = d0 * nr, cols = d2 * nc;
  std::vector<int64_t> np(cols + 1, 0), ni;
  std::vector<double> nx;
  for (int64_t c = 0; c < cols; ++c) {
    const int64_t oc = (c % d2) + d2 * (c / d2 + c0);
    for (int64_t q = p[oc]; q < p[oc + 1]; ++q) {
      const int64_t orow = i[q], bl = orow / d0;
      if (bl < r0 || bl > r1) continue;
      ni.push_back(orow % d0 + d0 * (bl - r0));
      nx.push_back(x[2 * q]);
      nx.push_back(x[2 * q + 1]);
    }
    np[c + 1] = static_cast<int64_t>(ni.size());
  }
  (void)wr;
  return tensor_from_csc(W.rules, std::move(idx), 2, rows, cols, np.data(), ni.data(), nx.data(), s);
}

TensorPtr bond_hamiltonian(Ctx& ctx, const BSTensor& W, bool odd) {
  CAM_REQUIRE(W.nidx() == 4, "bond Hamiltonian needs a 4-index MPO");
  const int64_t wl = W.idx[1].size, wr = W.idx[3].size;
  TensorPtr left, right;
  if (odd) {
    left = extract_mpo_block(ctx, W, wl - 1, wl - 1, 0, wr - 1);
    right = extract_mpo_block(ctx, W, 0, wl - 1, 0, 0);
  } else {
    CAM_REQUIRE(wl >= 3 && wr >= 3, "Even bond is empty");
    left = extract_mpo_block(ctx, W, wl - 1, wl - 1, 1, wr - 2);
    right = extract_mpo_block(ctx, W, 1, wl - 2, 0, 0);
  }
  TensorPtr t = contract(ctx, *left, false, *right, false, {{3, 1}});
  return reorder(ctx, *t, false, {0, 3, 2, 4, 1, 5}, 2);
}

TensorPtr bond_evolution_operator(Ctx& ctx, const BSTensor& H, double timestep) {
  const int N = H.nidx();
  CAM_REQUIRE(N >= 4 && H.numrows == 2, "bond operator needs [s1', s2', s1, s2, dummies...] with two row indices");
  for (int k = 4; k < N; ++k) CAM_REQUIRE(H.idx[k].size == 1, "bond operator: indices beyond the fourth must be dummies");
  const int64_t d0 = H.idx[0].size, d2 = H.idx[2].size;
  CAM_REQUIRE(H.rows() == H.cols(), "Hermitian matrix not square!");
  hipStream_t s = ctx.stream;
  const double eps = std::numeric_limits<double>::epsilon();

  // column groups (GetAllBlockColumns): outgoing charge of every column,
  // dummies included
  Charge qd = H.rules.zero();
  for (int k = 4; k < N; ++k) qd = H.idx[k].inward ? H.rules.sub(qd, H.idx[k].charges[0]) : H.rules.add(qd, H.idx[k].charges[0]);
  CAM_REQUIRE(qd == H.rules.zero(), "bond operator: dummy indices carry a net charge");
  const int64_t n = H.cols();
  std::map<Charge, int> gid_of;
  std::vector<int> gid(n);
  std::vector<int64_t> pos(n), gsize;
  for (int64_t c = 0; c < n; ++c) {
    Charge q = H.rules.zero();
    const int64_t i2 = c % d2, i3 = c / d2;
    q = H.idx[2].inward ? H.rules.sub(q, H.idx[2].charges[i2]) : H.rules.add(q, H.idx[2].charges[i2]);
    q = H.idx[3].inward ? H.rules.sub(q, H.idx[3].charges[i3]) : H.rules.add(q, H.idx[3].charges[i3]);
    auto it = gid_of.find(q);
    if (it == gid_of.end()) {
      it = gid_of.emplace(q, static_cast<int>(gsize.size())).first;
      gsize.push_back(0);
    }
    gid[c] = it->second;
    pos[c] = gsize[it->second]++;
  }
  const int ng = static_cast<int>(gsize.size());
  std::vector<int64_t> goff(ng + 1, 0);
  for (int g = 0; g < ng; ++g) {
    CAM_REQUIRE(gsize[g] < (int64_t(1) << 31), "bond operator block too large");
    goff[g + 1] = goff[g] + gsize[g] * gsize[g];
  }
  const int64_t gtot = goff[ng];

  // gather H into the dense group blocks
  std::vector<int64_t> mapA(gtot, -1);
  std::vector<int64_t> gp(N);
  for (const Block& b : H.blocks) {
    for (int64_t e = 0; e < b.size; ++e) {
      block_positions(H, b, e, gp.data());
      const int64_t R = gp[0] + d0 * gp[1], C = gp[2] + d2 * gp[3];
      CAM_REQUIRE(gid[R] == gid[C], "bond operator: row and column charges differ");
      const int g = gid[C];
      mapA[goff[g] + pos[R] + pos[C] * gsize[g]] = b.offset + e;
    }
  }
  DevBuf<int64_t> dmap(std::max<int64_t>(gtot, 1));
  dmap.upload(mapA.data(), gtot, s);
  DevBuf<cplx> A(std::max<int64_t>(gtot, 1)), B(std::max<int64_t>(gtot, 1)), Gd(std::max<int64_t>(gtot, 1));
  hipLaunchKernelGGL(map_gather_kernel, dim3(grid_of(gtot)), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(H.data.get()), dmap.get(), reinterpret_cast<double2*>(A.get()),
                     gtot);
  HIP_CHECK(hipGetLastError());

  rocblas_handle h = static_cast<rocblas_handle>(ctx.blas_handle());
  int64_t gmax = 1;
  for (int g = 0; g < ng; ++g) gmax = std::max(gmax, gsize[g]);
  DevBuf<double> lam(gmax);
  for (int g = 0; g < ng; ++g) {
    const int64_t m = gsize[g];
    auto* Ag = reinterpret_cast<rocblas_double_complex*>(A.get() + goff[g]);
    hermitian_eig(ctx, reinterpret_cast<cplx*>(Ag), m, lam.get());
    // exp(factor * lambda) with factor = (0, -timestep)
    hipLaunchKernelGGL(exp_columns_kernel, dim3(grid_of(m * m)), dim3(256), 0, s,
                       reinterpret_cast<double2*>(A.get() + goff[g]), reinterpret_cast<double2*>(B.get() + goff[g]),
                       lam.get(), m, 0.0, -timestep, eps / static_cast<double>(m));
    HIP_CHECK(hipGetLastError());
    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
    const rocblas_status st = rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_conjugate_transpose, static_cast<rocblas_int>(m),
                       static_cast<rocblas_int>(m), static_cast<rocblas_int>(m), &one,
                       reinterpret_cast<const rocblas_double_complex*>(B.get() + goff[g]), static_cast<rocblas_int>(m),
                       Ag, static_cast<rocblas_int>(m), &zero,
                       reinterpret_cast<rocblas_double_complex*>(Gd.get() + goff[g]), static_cast<rocblas_int>(m));
    CAM_REQUIRE(st == rocblas_status_success, "rocblas_zgemm failed");
  }

  // G: the first four indices, every zero-flux block
  auto G = std::make_shared<BSTensor>();
  G->rules = H.rules;
  G->idx.assign(H.idx.begin(), H.idx.begin() + 4);
  G->numrows = 2;
  for (auto& sec : G->allowed_sector_tuples()) {
    Block b;
    b.sec = sec;
    b.dims.resize(4);
    for (int k = 0; k < 4; ++k) b.dims[k] = G->idx[k].sec_size[sec[k]];
    G->blocks.push_back(b);
  }
  G->finalize_layout();
  std::vector<int64_t> mapG(std::max<int64_t>(G->total, 1), -1);
  for (const Block& b : G->blocks) {
    for (int64_t e = 0; e < b.size; ++e) {
      block_positions(*G, b, e, gp.data());
      const int64_t R = gp[0] + d0 * gp[1], C = gp[2] + d2 * gp[3];
      const int g = gid[C];
      if (gid[R] != g) continue;
      mapG[b.offset + e] = goff[g] + pos[R] + pos[C] * gsize[g];
    }
  }
  DevBuf<int64_t> dmapG(std::max<int64_t>(G->total, 1));
  dmapG.upload(mapG.data(), G->total, s);
  if (G->total)
    hipLaunchKernelGGL(map_gather_kernel, dim3(grid_of(G->total)), dim3(256), 0, s,
                       reinterpret_cast<const double2*>(Gd.get()), dmapG.get(),
                       reinterpret_cast<double2*>(G->data.get()), G->total);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(s));
  return G;
}

}  // namespace cam
