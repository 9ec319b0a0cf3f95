// Complex Schur decomposition of small dense matrices on the host, for the
// projected matrices (at most 64 x 64) of the Krylov-Schur eigensolver.
#pragma once
#include <complex>
#include <vector>

namespace cam {

using zcplx = std::complex<double>;

// A (n x n, column-major) = Z T Z^H with T upper triangular and Z unitary;
// A is overwritten by T.  Householder reduction to Hessenberg form, then
// single-shift QR sweeps with Wilkinson shifts (an exceptional shift every
// 11th sweep on one eigenvalue).  Returns false when an eigenvalue has not
// converged after 30 n sweeps.
bool complex_schur(int n, std::vector<zcplx>& A, std::vector<zcplx>& Z);

// Exchange the adjacent eigenvalues T(k,k) and T(k+1,k+1) by a Givens
// rotation, updating T and Z (the rotation of LAPACK ztrexc).
void schur_swap(int n, std::vector<zcplx>& T, std::vector<zcplx>& Z, int k);

// Bring the `count` eigenvalues of largest magnitude to the leading
// positions, in descending magnitude.
void schur_order_by_magnitude(int n, std::vector<zcplx>& T, std::vector<zcplx>& Z, int count);

}  // namespace cam
