// Orthogonalisation of a unit cell by its transfer-matrix fixed points, and
// the energy per vertex of a two-vertex cell, on the device.
#pragma once
#include <functional>
#include <vector>

#include "solvers.hpp"

namespace cam {

// The left / right transfer maps of a unit cell [A_0 .. A_{L-1}] with itself
// (TransferMatrixComponents with Bra = Ket, measurement.cpp:405-449):
//   left:  v -> sum_s A_s^H v A_s through the cell (LeftComponentsMultiply,
//          582-602), v on (bra inward of A_0) x (ket inward of A_0);
//   right: v -> sum_s conj(A_s) v A_s^T backwards (RightComponentsMultiply,
//          603-629), v on (bra outward of A_{L-1}) x (ket outward).
// A vector holds the charge-diagonal blocks v[r, c] (equal charges of r and
// c, target = identity) of the boundary index, block s an n_s x n_s
// column-major array (rows bra), blocks in sector order, back to back: the
// layout of a tensor with indices (inward q, outward q) and blocks (s, s).
// Each site is one permuted copy of A (in, phys, out); a site step is two
// grouped MFMA launches (v . A, then A^H . (v A) for the left map).
struct TransferOp {
  Ctx* ctx;
  std::vector<TensorPtr> P;                 // per site: A reordered to (in, phys, out)
  struct Bond {
    std::vector<int64_t> n, off;            // per sector: size, block offset
    int64_t total = 0;
  };
  std::vector<Bond> bonds;                  // bonds[i] = inward index of site i; bonds[L] = outward of the last
  const QIndex* boundary = nullptr;         // inward index of A_0
  GemmPlan plan[2][2][8];                   // [side][gemm][site]
  std::vector<std::pair<int64_t, int64_t>> empty[2][8];   // output ranges no block writes
  DevBuf<cplx> wbuf, mid[2];
  bool tuned[2] = {false, false};
  double flops[2] = {0.0, 0.0};             // per matvec
  int64_t n() const { return bonds.front().total; }

  TransferOp(Ctx& c, const std::vector<TensorPtr>& cell);
  void apply(int side, const cplx* in, cplx* out);
};

// Largest-magnitude eigenpair of a linear map on length-n device vectors by
// Krylov-Schur (thick restart of m-step Arnoldi keeping the m/2 Schur
// vectors of largest magnitude), Arnoldi vectors orthogonalised by two
// classical Gram-Schmidt passes.  x holds the start vector and receives the
// unit eigenvector.  Converged when the Ritz residual |beta e_m^T z| is
// <= tol |theta| (ARPACK's criterion with nev = 1, ncv = m).  With
// `project`, the iteration runs on a real subspace invariant under the map:
// project(v) maps a product onto it and the Gram-Schmidt coefficients are
// taken real (the start vector must lie in the subspace).
struct LeadingEig {
  cplx value{0.0, 0.0};
  int matvecs = 0, restarts = 0;
  double residual = 0.0;
  bool converged = false;
};
LeadingEig leading_eigenpair(Ctx& ctx, const std::function<void(const cplx*, cplx*)>& mv, int64_t n, cplx* x,
                             int m, int max_restarts, double tol,
                             const std::function<void(cplx*)>& project = nullptr);

// LeftED / RightED (measurement.cpp:451-580) with NUMEVALS = 1 and
// LARGESTMAGNITUDE: the leading eigenvalue, and (when M != nullptr) the
// Hermitian completion v + v^H of its eigenvector rotated to a real positive
// trace, as a tensor with indices (inward q, outward q), q the boundary.
// ncv <= 0 takes ARPACK's 20.
LeadingEig transfer_leading(Ctx& ctx, const std::vector<TensorPtr>& cell, int side, double tol, TensorPtr* M,
                            int ncv = 0);
// One application of the left (0) or right (1) map to a host vector of the
// allowed entries in the reference's order (LeftComponentsMultiply /
// RightComponentsMultiply: c outer, r inner, positions r + n c).
void transfer_apply(Ctx& ctx, const std::vector<TensorPtr>& cell, int side, const cplx* in, cplx* out);

struct OrthoInfo {
  LeadingEig left, right;
  double entropy = 0.0;
  int64_t kept_left = 0, kept_right = 0;
  double seconds = 0.0;
};

// UnitCell(MPSDecomposition, PreviousLambda) (UnitCell.hpp:42-48): the
// second matrix diag(values) . Vd . diag(previous)^-1 reordered (phys, in, out).
TensorPtr unit_cell_second(Ctx& ctx, const BSTensor& Vd, const std::vector<double>& values,
                           const std::vector<double>& previous);

// Orthogonalise (measurement.cpp:180-246) of the two-vertex cell [A0, A1]
// whose second lambda is lambda1.  Throws on a complex leading eigenvalue or
// a SqrtDR failure (the reference returns an empty cell there).
void orthogonalise(Ctx& ctx, const TensorPtr& A0, const TensorPtr& A1, TensorPtr& A0n, TensorPtr& A1n,
                   std::vector<double>& lambda0n, OrthoInfo& info);

// OrthogonaliseInversionSymmetric (measurement.cpp:131-178): only the left
// fixed point is solved and serves on the right boundary too; the boundary
// lambda lambda0 (on the last matrix's outward index) enters between the
// factors, SVD of X . diag(lambda0) . Y.
void orthogonalise_inversion_symmetric(Ctx& ctx, const TensorPtr& A0, const TensorPtr& A1,
                                       const std::vector<double>& lambda0, TensorPtr& A0n, TensorPtr& A1n,
                                       std::vector<double>& lambda0n, OrthoInfo& info);

// Overlap(bra, ket) (measurement.cpp:282-288): the largest-magnitude
// eigenvalue of TransferMatrixComponents(bra, ket, 0, State(rules)), the
// left map v -> sum_s B_s^H v A_s on the charge-paired blocks of (bra
// inward) x (ket inward).
cplx overlap(Ctx& ctx, const std::vector<TensorPtr>& bra, const std::vector<TensorPtr>& ket);

// OneVertexMeasurement (measurement.cpp:317-324): the MPO operator Op on the
// second vertex of the cell, closed with lambda0^2.
cplx one_vertex_measurement(Ctx& ctx, const BSTensor& Op, const TensorPtr& A0, const TensorPtr& A1,
                            const std::vector<double>& lambda0);

// iTwoVertexEnergy (measurement.cpp:371-397) of the orthogonalised cell
// [A0, A1] with boundary lambda lambda0, from the Hamiltonian MPO W.
// terms: on-vertex, bond inside the cell, bond to the next cell.
cplx i_two_vertex_energy(Ctx& ctx, const BSTensor& W, const TensorPtr& A0, const TensorPtr& A1,
                         const std::vector<double>& lambda0, cplx terms[3]);

}  // namespace cam
