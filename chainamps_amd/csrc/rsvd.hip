// Leading singular triplets of a large dense complex block by randomized
// subspace iteration (Halko, Martinsson & Tropp, SIAM Rev. 53 (2011) 217,
// algorithms 4.4 and 5.1), all on the device with rocBLAS/rocSOLVER:
//   Y = A Omega (n x k Gaussian-like start), then q rounds of
//   Q = orth(Y), Z = orth(A^H Q), Y = A Z;   Q = orth(Y);
//   A^H Q = Q2 R  (QR),  R^H = Ur S Vr^H  (k x k SVD, Jordan-Wielandt),
//   U = Q Ur,  V^H = Vr^H Q2^H.
// The singular values that matter are accepted only when their residuals
// |A v_j - s_j u_j| are below `tol` (A^H u_j = s_j v_j holds by construction,
// so this bounds the error of s_j); otherwise more rounds run, and the caller falls back to the
// full decomposition when the budget is spent.  The work is GEMM-shaped
// (k columns at a time), O(m n k) per round instead of O(m n min(m, n)).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>

#include "solvers.hpp"

namespace cam {

namespace {

// out (k x k) = R^H with R the upper triangle of the k x k top of `qr` (ld)
__global__ void upper_conj_transpose_kernel(const double2* __restrict__ qr, int64_t ld, int64_t k,
                                            double2* __restrict__ out) {
  const int64_t tot = k * k;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < tot;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = e % k, j = e / k;   // out(i, j) = conj(R(j, i)), R(j, i) = 0 for j > i
    double2 v = make_double2(0.0, 0.0);
    if (j <= i) {
      const double2 r = qr[j + i * ld];
      v = make_double2(r.x, -r.y);
    }
    out[e] = v;
  }
}

// E(:, j) -= s_j U(:, j), then res[j] = |E(:, j)|^2 (one block per column;
// E and U are m x ncheck / m x k, column-major)
__global__ void ritz_residual_kernel(double2* __restrict__ E, int64_t m, const double2* __restrict__ U,
                                     const double* __restrict__ s, double* __restrict__ res) {
  __shared__ double red[256];
  const int64_t j = blockIdx.x;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < m; i += blockDim.x) {
    const double2 u = U[i + j * m];
    double2 e = E[i + j * m];
    e.x -= u.x * s[j];
    e.y -= u.y * s[j];
    acc += e.x * e.x + e.y * e.y;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) res[j] = red[0];
}

int grid_of(int64_t n) { return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(8192, (n + 255) / 256))); }

const rocblas_double_complex kOne{1.0, 0.0}, kZero{0.0, 0.0};

rocblas_double_complex* Z(cplx* p) { return reinterpret_cast<rocblas_double_complex*>(p); }

void gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int64_t m, int64_t n, int64_t k,
          const cplx* A, int64_t lda, const cplx* B, int64_t ldb, cplx* C, int64_t ldc) {
  const rocblas_status st =
      rocblas_zgemm(h, ta, tb, static_cast<rocblas_int>(m), static_cast<rocblas_int>(n),
                    static_cast<rocblas_int>(k), &kOne, reinterpret_cast<const rocblas_double_complex*>(A),
                    static_cast<rocblas_int>(lda), reinterpret_cast<const rocblas_double_complex*>(B),
                    static_cast<rocblas_int>(ldb), &kZero, Z(C), static_cast<rocblas_int>(ldc));
  CAM_REQUIRE(st == rocblas_status_success, "rocblas_zgemm failed");
}

// X (m x k, ld m) <- orthonormal basis of its columns (Householder QR, explicit Q)
void orthonormalise(rocblas_handle h, cplx* X, int64_t m, int64_t k, DevBuf<cplx>& tau) {
  tau.ensure(std::max<int64_t>(k, 1));
  rocblas_status st = rocsolver_zgeqrf(h, static_cast<rocblas_int>(m), static_cast<rocblas_int>(k), Z(X),
                                       static_cast<rocblas_int>(m), Z(tau.get()));
  CAM_REQUIRE(st == rocblas_status_success, "rocsolver_zgeqrf failed");
  st = rocsolver_zungqr(h, static_cast<rocblas_int>(m), static_cast<rocblas_int>(k), static_cast<rocblas_int>(k), Z(X),
                        static_cast<rocblas_int>(m), Z(tau.get()));
  CAM_REQUIRE(st == rocblas_status_success, "rocsolver_zungqr failed");
}

}  // namespace

bool truncated_svd(Ctx& ctx, const cplx* A, int64_t m, int64_t n, int64_t k, int64_t ncheck, double tol,
                   int max_rounds, std::vector<double>& S, DevBuf<cplx>& U, DevBuf<cplx>& Vh, int* rounds) {
  CAM_REQUIRE(k >= 1 && k <= std::min(m, n), "truncated_svd: k out of range");
  CAM_REQUIRE(m < (int64_t(1) << 31) && n < (int64_t(1) << 31), "truncated_svd: block too large");
  hipStream_t s = ctx.stream;
  rocblas_handle h = static_cast<rocblas_handle>(ctx.blas_handle());
  ncheck = std::min(ncheck, k);
  DevBuf<cplx> Y(m * k), W(n * k), R(k * k), Ur(k * k), Vrh(k * k), E(m * std::max<int64_t>(ncheck, 1)), tau;
  DevBuf<double> dS(k), dres(std::max<int64_t>(ncheck, 1));
  launch_random_fill(W.get(), n * k, 0x5eed0000ull + static_cast<uint64_t>(m * 31 + n), s);
  gemm(h, rocblas_operation_none, rocblas_operation_none, m, k, n, A, m, W.get(), n, Y.get(), m);
  int q = 0;
  for (;;) {
    // two subspace rounds between residual checks
    for (int r = 0; r < 2; ++r, ++q) {
      orthonormalise(h, Y.get(), m, k, tau);
      gemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, n, k, m, A, m, Y.get(), m, W.get(), n);
      orthonormalise(h, W.get(), n, k, tau);
      gemm(h, rocblas_operation_none, rocblas_operation_none, m, k, n, A, m, W.get(), n, Y.get(), m);
    }
    // Rayleigh-Ritz on span(Y)
    DevBuf<cplx> Q(m * k);
    HIP_CHECK(hipMemcpyAsync(Q.get(), Y.get(), m * k * sizeof(cplx), hipMemcpyDeviceToDevice, s));
    orthonormalise(h, Q.get(), m, k, tau);
    gemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, n, k, m, A, m, Q.get(), m, W.get(), n);
    tau.ensure(k);
    rocblas_status st = rocsolver_zgeqrf(h, static_cast<rocblas_int>(n), static_cast<rocblas_int>(k), Z(W.get()),
                                         static_cast<rocblas_int>(n), Z(tau.get()));
    CAM_REQUIRE(st == rocblas_status_success, "rocsolver_zgeqrf failed");
    hipLaunchKernelGGL(upper_conj_transpose_kernel, dim3(grid_of(k * k)), dim3(256), 0, s,
                       reinterpret_cast<const double2*>(W.get()), n, k, reinterpret_cast<double2*>(R.get()));
    HIP_CHECK(hipGetLastError());
    st = rocsolver_zungqr(h, static_cast<rocblas_int>(n), static_cast<rocblas_int>(k), static_cast<rocblas_int>(k),
                          Z(W.get()), static_cast<rocblas_int>(n), Z(tau.get()));
    CAM_REQUIRE(st == rocblas_status_success, "rocsolver_zungqr failed");
    // k x k core R^H = Ur S Vr^H through the Hermitian eigenproblem of
    // [[0, R^H], [R, 0]] (divide and conquer; no squared condition)
    std::vector<double> core;
    jordan_wielandt_svd(ctx, R.get(), k, k, k, k, core, Ur.get(), Vrh.get());
    dS.upload(core.data(), k, s);
    U.alloc(m * k);
    Vh.alloc(k * n);
    gemm(h, rocblas_operation_none, rocblas_operation_none, m, k, k, Q.get(), m, Ur.get(), k, U.get(), m);
    gemm(h, rocblas_operation_none, rocblas_operation_conjugate_transpose, k, n, k, Vrh.get(), k, W.get(), n,
         Vh.get(), k);
    // residuals |A v_j - s_j u_j| of the leading ncheck triplets: by
    // construction A^H u_j = s_j v_j exactly, so |s_j - sigma| <= this
    // residual for some singular value sigma of A
    gemm(h, rocblas_operation_none, rocblas_operation_conjugate_transpose, m, ncheck, n, A, m, Vh.get(), k, E.get(),
         m);
    hipLaunchKernelGGL(ritz_residual_kernel, dim3(ncheck), dim3(256), 0, s, reinterpret_cast<double2*>(E.get()), m,
                       reinterpret_cast<const double2*>(U.get()), dS.get(), dres.get());
    HIP_CHECK(hipGetLastError());
    S = core;
    std::vector<double> res(ncheck);
    HIP_CHECK(hipMemcpyAsync(res.data(), dres.get(), ncheck * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    double worst = 0.0;
    for (double r : res) worst = std::max(worst, std::sqrt(r));
    if (rounds) *rounds = q;
    if (worst <= tol) return true;
    if (q >= max_rounds) return false;
    // continue iterating from the current subspace
    HIP_CHECK(hipMemcpyAsync(Y.get(), U.get(), m * k * sizeof(cplx), hipMemcpyDeviceToDevice, s));
  }
}

}  // namespace cam
