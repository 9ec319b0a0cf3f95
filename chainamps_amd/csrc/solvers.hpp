// Eigensolver, SVD truncation, iTEBD gate and adjoint on the device.
#pragma once
#include <vector>

#include "ctx.hpp"
#include "heff.hpp"

namespace cam {

// Lowest eigenpair of H_eff (replaces arpack_eigs at TwoVertexComponents::HED).
void heff_lowest(Heff& H, const cplx* guess_ref, bool converge, double tol, int ncv, bool mixed,
                 double* eval, cplx* evec_ref, double* info);

// Block-layout core: x (device, n) holds the start vector when have_start
// and receives the gauge-fixed eigenvector; returns the eigenvalue.
double lanczos_lowest(Heff& H, cplx* x, bool have_start, bool converge, double tol, int ncv,
                      bool mixed, double* info);
// Dense branch for n <= SPARSE_THRESHOLD (sparse_interface.hpp:16) and the
// dispatcher that picks it, as TwoVertexComponents::HED does.
constexpr int64_t kSparseThreshold = 1000;
double dense_lowest(Heff& H, cplx* x, double* info);
double eigs_lowest(Heff& H, cplx* x, bool have_start, bool converge, double tol, int ncv, bool mixed,
                   double* info);
void gauge_fix(cplx* x, int64_t n, hipStream_t s);
// entropy of singular values: sum over nonzero s of -s^2 log s^2
// (sparse_interface.cpp:73-83)
double entropy(const std::vector<double>& values);
// contract with a diagonal on index k (values per position of that index)
void scale_along_index(Ctx& ctx, BSTensor& T, int k, const std::vector<double>& lam);

// Leading k singular triplets of a dense m x n block (column-major, device)
// by randomized subspace iteration; true when the Ritz residuals of the
// first ncheck triplets are <= tol within max_rounds rounds.  U is m x k,
// Vh is k x n, S descending.  last_res (if given) receives the residual
// |A v_k - s_k u_k| of the last computed triplet at acceptance (0 when
// ncheck == k, whose residuals are all within tol).
// With naccept (0 < naccept <= ncheck) only the first naccept residuals
// must be <= tol; res (if given) receives all ncheck residuals of the
// accepted result.
bool truncated_svd(Ctx& ctx, const cplx* A, int64_t m, int64_t n, int64_t k, int64_t ncheck, double tol,
                   int max_rounds, std::vector<double>& S, DevBuf<cplx>& U, DevBuf<cplx>& Vh, int* rounds,
                   double* last_res = nullptr, int64_t naccept = -1, std::vector<double>* res = nullptr);

// Hermitian eigendecomposition of the n x n device matrix A (column-major):
// on return A holds the eigenvectors, w (device, n) the eigenvalues in
// ascending order.  Divide and conquer, every eigenpair checked, QR
// iteration as the fallback.
void hermitian_eig(Ctx& ctx, cplx* A, int64_t n, double* w);
// Leading k singular triplets of the m x n device matrix B (leading
// dimension ldb) from the eigenpairs of [[0, B], [B^H, 0]]: S descending
// (host), U m x k, Vh k x n (device, column-major).
void jordan_wielandt_svd(Ctx& ctx, const cplx* B, int64_t ldb, int64_t m, int64_t n, int64_t k,
                         std::vector<double>& S, cplx* U, cplx* Vh);

// MPX_matrix::SVD with the reference truncation rule.
void svd_truncate(Ctx& ctx, const BSTensor& theta, int64_t D, double max_trunc, TensorPtr& U,
                  TensorPtr& Vd, std::vector<double>& values, double& discarded);

// The singular values MPX_matrix::SVD would keep with no truncation (D = 0,
// no truncation weight), without computing the singular vectors.
void singular_values(Ctx& ctx, const BSTensor& theta, std::vector<double>& values);

// iTEBD::apply_and_decompose.
void tebd_gate(Ctx& ctx, const BSTensor& G, const BSTensor& A0, const BSTensor& A1,
               const std::vector<double>& lambda0, int64_t D, double minS, TensorPtr& A0n,
               TensorPtr& A1n, std::vector<double>& lambda1, double& truncation);

// Gate construction: ExtractMPOBlock, odd/even bond Hamiltonians,
// exp(-i timestep H) per charge block.
TensorPtr extract_mpo_block(Ctx& ctx, const BSTensor& W, int64_t r0, int64_t r1, int64_t c0, int64_t c1);
TensorPtr bond_hamiltonian(Ctx& ctx, const BSTensor& W, bool odd);
TensorPtr bond_evolution_operator(Ctx& ctx, const BSTensor& H, double timestep);

// dE/dconj(psi) and dE/dW for E = <psi|H|psi>/<psi|psi>.
void heff_vjp(Heff& H, const cplx* psi_ref, cplx* dpsi_ref, cplx* dW, int64_t nnzW);

}  // namespace cam
