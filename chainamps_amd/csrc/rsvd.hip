// Leading singular triplets of a large dense complex block by randomized
// subspace iteration (Halko, Martinsson & Tropp, SIAM Rev. 53 (2011) 217,
// algorithms 4.4 and 5.1), all on the device with rocBLAS/rocSOLVER:
//   Y = A Omega (n x k Gaussian-like start), then q rounds of
//   Q = orth(Y), Z = orth(A^H Q), Y = A Z;   Q = orth(Y);
//   A^H Q = Q2 R  (Cholesky QR),  R^H = Ur S Vr^H  (k x k SVD, Jordan-Wielandt),
//   U = Q Ur,  V^H = Vr^H Q2^H.
// The singular values that matter are accepted only when their residuals
// |A v_j - s_j u_j| are below `tol` (A^H u_j = s_j v_j holds by construction,
// so this bounds the error of s_j); otherwise two more rounds run, and the caller falls back to the
// full decomposition when the budget is spent.  The work is GEMM-shaped
// (k columns at a time), O(m n k) per round instead of O(m n min(m, n)).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>

#include "solvers.hpp"

namespace cam {

namespace {

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
void householder_orthonormalise(rocblas_handle h, cplx* X, int64_t m, int64_t k, DevBuf<cplx>& tau) {
  tau.ensure(std::max<int64_t>(k, 1));
  rocblas_status st = rocsolver_zgeqrf(h, static_cast<rocblas_int>(m), static_cast<rocblas_int>(k), Z(X),
                                       static_cast<rocblas_int>(m), Z(tau.get()));
  CAM_REQUIRE(st == rocblas_status_success, "rocsolver_zgeqrf failed");
  st = rocsolver_zungqr(h, static_cast<rocblas_int>(m), static_cast<rocblas_int>(k), static_cast<rocblas_int>(k), Z(X),
                        static_cast<rocblas_int>(m), Z(tau.get()));
  CAM_REQUIRE(st == rocblas_status_success, "rocsolver_zungqr failed");
}

// G (k x k, upper triangle) += c * trace(G) on the diagonal (one workgroup)
__global__ void shift_diagonal_kernel(double2* __restrict__ G, int64_t k, double c) {
  __shared__ double red[256];
  double t = 0.0;
  for (int64_t i = threadIdx.x; i < k; i += blockDim.x) t += G[i + i * k].x;
  red[threadIdx.x] = t;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double sh = c * red[0];
  for (int64_t i = threadIdx.x; i < k; i += blockDim.x) G[i + i * k].x += sh;
}

// out[0] = max |G_ij - delta_ij| over a k x k matrix (one workgroup)
__global__ void identity_deviation_kernel(const double2* __restrict__ G, int64_t k, double* __restrict__ out) {
  __shared__ double red[256];
  double m = 0.0;
  for (int64_t e = threadIdx.x; e < k * k; e += blockDim.x) {
    const double2 g = G[e];
    const double re = g.x - ((e % k) == (e / k) ? 1.0 : 0.0);
    m = fmax(m, fmax(fabs(re), fabs(g.y)));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// True when the columns of X are orthonormal to 1e-12 (max |X^H X - I|).
bool orthonormal(Ctx& ctx, const cplx* X, int64_t m, int64_t k, cplx* G) {
  rocblas_handle h = static_cast<rocblas_handle>(ctx.blas_handle());
  hipStream_t s = ctx.stream;
  gemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, k, k, m, X, m, X, m, G, k);
  DevBuf<double> dev(1);
  hipLaunchKernelGGL(identity_deviation_kernel, dim3(1), dim3(256), 0, s, reinterpret_cast<const double2*>(G), k,
                     dev.get());
  HIP_CHECK(hipGetLastError());
  double v = 1.0;
  HIP_CHECK(hipMemcpyAsync(&v, dev.get(), sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return v <= 1e-12;
}

// ratio[0] = min_i |R_ii| / max_i |R_ii| of a k x k Cholesky factor (one
// workgroup); zero when the factorization reported a breakdown
__global__ void pivot_ratio_kernel(const double2* __restrict__ R, int64_t k, const int* __restrict__ info,
                                   double* __restrict__ ratio) {
  __shared__ double lo[256], hi[256];
  double a = INFINITY, b = 0.0;
  for (int64_t i = threadIdx.x; i < k; i += blockDim.x) {
    const double v = fabs(R[i + i * k].x);
    a = fmin(a, v);
    b = fmax(b, v);
  }
  lo[threadIdx.x] = a;
  hi[threadIdx.x] = b;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + o]);
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) ratio[0] = (info[0] != 0 || !(hi[0] > 0.0)) ? 0.0 : lo[0] / hi[0];
}

// One Cholesky QR pass: G = X^H X (+ shift), R = chol(G), X <- X R^{-1}.
// False when the factorization breaks down or, with min_pivot > 0, when
// min |R_ii| < min_pivot max |R_ii| (X too close to rank deficient for
// one pass); X is then unchanged.
bool cholqr_pass(Ctx& ctx, cplx* X, int64_t m, int64_t k, cplx* G, int* dinfo, double shift,
                 double min_pivot = 0.0, double* dratio = nullptr) {
  rocblas_handle h = static_cast<rocblas_handle>(ctx.blas_handle());
  hipStream_t s = ctx.stream;
  // the Gram matrix with zgemm: rocBLAS zherk takes 13 ms where zgemm takes
  // 1.8 ms for 16000 x 750 on MI355X
  gemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, k, k, m, X, m, X, m, G, k);
  rocblas_status st;
  if (shift > 0.0) {
    hipLaunchKernelGGL(shift_diagonal_kernel, dim3(1), dim3(256), 0, s, reinterpret_cast<double2*>(G), k, shift);
    HIP_CHECK(hipGetLastError());
  }
  st = rocsolver_zpotrf(h, rocblas_fill_upper, static_cast<rocblas_int>(k), Z(G), static_cast<rocblas_int>(k), dinfo);
  CAM_REQUIRE(st == rocblas_status_success, "rocsolver_zpotrf failed");
  if (min_pivot > 0.0) {
    hipLaunchKernelGGL(pivot_ratio_kernel, dim3(1), dim3(256), 0, s, reinterpret_cast<const double2*>(G), k, dinfo,
                       dratio);
    HIP_CHECK(hipGetLastError());
    double ratio = 0.0;
    HIP_CHECK(hipMemcpyAsync(&ratio, dratio, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (!(ratio >= min_pivot)) return false;
  } else {
    int info = 0;
    HIP_CHECK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (info != 0) return false;
  }
  st = rocblas_ztrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit,
                     static_cast<rocblas_int>(m), static_cast<rocblas_int>(k), &kOne, Z(G),
                     static_cast<rocblas_int>(k), Z(X), static_cast<rocblas_int>(m));
  CAM_REQUIRE(st == rocblas_status_success, "rocblas_ztrsm failed");
  return true;
}

// X (m x k) <- orthonormal basis of its columns by Cholesky QR2, or, when
// its Cholesky factorization breaks down, shifted Cholesky QR3 (Fukaya,
// Kannan, Nakatsukasa, Yamamoto, Yanagisawa, SIAM J. Sci. Comput. 42 (2020)
// A477): one pass with the shift 11 (m k + k (k + 1)) u |X|_F^2, which keeps
// the factorization defined for any condition below 1/u, then two plain
// passes; GEMM-shaped passes instead of a Householder QR.  Householder QR
// takes over if a factorization still breaks down (numerically
// rank-deficient X).
// X[i] += scale * (a, b), a and b uniform in [-1, 1), scale = rel * the RMS
// of X's elements (|X|^2 from nrm2[0].x), from a counter-based stream
__global__ void perturb_kernel(double2* __restrict__ X, int64_t n, uint64_t seed, const double2* __restrict__ nrm2,
                               double rel) {
  const double scale = rel * sqrt(fmax(nrm2[0].x, 0.0) / static_cast<double>(n));
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * static_cast<uint64_t>(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double a = (z >> 11) * (1.0 / 9007199254740992.0);
    z = z * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    z ^= z >> 29;
    const double b = (z >> 11) * (1.0 / 9007199254740992.0);
    double2 x = X[i];
    x.x += scale * (2.0 * a - 1.0);
    x.y += scale * (2.0 * b - 1.0);
    X[i] = x;
  }
}

void orthonormalise(Ctx& ctx, cplx* X, int64_t m, int64_t k, DevBuf<cplx>& G, DevBuf<int>& dinfo,
                    DevBuf<cplx>& tau) {
  G.ensure(static_cast<size_t>(k) * k);
  dinfo.ensure(1);
  const double u = 0.5 * std::numeric_limits<double>::epsilon();
  const double c = 11.0 * (double(m) * double(k) + double(k) * double(k + 1)) * u;
  // plain Cholesky QR2 (enough for conditions up to about 1/sqrt(u)), then
  // shifted Cholesky QR3, then Householder; each result is accepted only
  // when max |X^H X - I| <= 1e-12
  if (cholqr_pass(ctx, X, m, k, G.get(), dinfo.get(), 0.0) && cholqr_pass(ctx, X, m, k, G.get(), dinfo.get(), 0.0) &&
      orthonormal(ctx, X, m, k, G.get()))
    return;
  if (cholqr_pass(ctx, X, m, k, G.get(), dinfo.get(), c) && cholqr_pass(ctx, X, m, k, G.get(), dinfo.get(), 0.0) &&
      cholqr_pass(ctx, X, m, k, G.get(), dinfo.get(), 0.0) && orthonormal(ctx, X, m, k, G.get()))
    return;
  // Numerically rank-deficient X (the block itself has rank below k, as in
  // the first iTEBD steps from a product state): complete it with random
  // directions of 1e-14 of its element RMS.  The span of its non-negligible
  // directions moves by about 1e-14 relative (below the acceptance
  // tolerance), the rest becomes a random orthonormal completion, and the
  // condition stays within the reach of shifted Cholesky QR3.  Householder
  // QR (rocSOLVER, panel-by-panel: about 1 s at 34000 x 1536) is the last
  // resort.
  {
    hipStream_t s = ctx.stream;
    const int64_t n = m * k;
    DevBuf<cplx> part(static_cast<size_t>(multi_dot_blocks(n)) * 64), nrm(64);
    launch_multi_dot(X, n, 1, X, n, part.get(), nrm.get(), s);
    static std::atomic<uint64_t> counter{0};
    const uint64_t seed = 0x9e3779b97f4a7c15ull * (counter.fetch_add(1) + 1) ^ static_cast<uint64_t>(n);
    hipLaunchKernelGGL(perturb_kernel, dim3(grid_of(n)), dim3(256), 0, s, reinterpret_cast<double2*>(X), n, seed,
                       reinterpret_cast<const double2*>(nrm.get()), 1e-14);
    HIP_CHECK(hipGetLastError());
    if (cholqr_pass(ctx, X, m, k, G.get(), dinfo.get(), c) && cholqr_pass(ctx, X, m, k, G.get(), dinfo.get(), 0.0) &&
        cholqr_pass(ctx, X, m, k, G.get(), dinfo.get(), 0.0) && orthonormal(ctx, X, m, k, G.get()))
      return;
  }
  householder_orthonormalise(static_cast<rocblas_handle>(ctx.blas_handle()), X, m, k, tau);
}

}  // namespace

bool truncated_svd(Ctx& ctx, const cplx* A, int64_t m, int64_t n, int64_t k, int64_t ncheck, double tol,
                   int max_rounds, std::vector<double>& S, DevBuf<cplx>& U, DevBuf<cplx>& Vh, int* rounds,
                   double* last_res, int64_t naccept, std::vector<double>* res_out) {
  CAM_REQUIRE(k >= 1 && k <= std::min(m, n), "truncated_svd: k out of range");
  CAM_REQUIRE(m < (int64_t(1) << 31) && n < (int64_t(1) << 31), "truncated_svd: block too large");
  hipStream_t s = ctx.stream;
  rocblas_handle h = static_cast<rocblas_handle>(ctx.blas_handle());
  ncheck = std::min(ncheck, k);
  naccept = (naccept <= 0) ? ncheck : std::min(naccept, ncheck);
  DevBuf<cplx> Y(m * k), W(n * k), Wc(n * k), R(k * k), Ur(k * k), Vrh(k * k), E(m * std::max<int64_t>(ncheck, 1)),
      tau, G;
  DevBuf<int> dinfo;
  DevBuf<double> dS(k), dres(std::max<int64_t>(ncheck, 1));
  launch_random_fill(W.get(), n * k, 0x5eed0000ull + static_cast<uint64_t>(m * 31 + n), s);
  gemm(h, rocblas_operation_none, rocblas_operation_none, m, k, n, A, m, W.get(), n, Y.get(), m);
  int q = 0;
  // The first residual check (a Rayleigh-Ritz projection, a k x k core and
  // the residuals, about 1.5 rounds' worth) waits for the rounds the last
  // accepted block of this shape on this context needed, since the blocks
  // of consecutive iDMRG / iTEBD steps converge alike (the hint only moves
  // the check; acceptance is always the residual test below).
  // (the number of checked triplets sets the pace: the slowest of them
  // converges at the rate (s_{k+1} / s_ncheck)^2 per round, so the hints
  // are kept per ncheck as well)
  const std::array<int64_t, 4> shape{m, n, k, naccept};
  const int first_default = (m * n > (int64_t(1) << 24)) ? 4 : 2;
  int first = first_default;
  {
    auto it = ctx.svd_round_hint.find(shape);
    if (it != ctx.svd_round_hint.end()) first = std::max(first, std::min(it->second[0], max_rounds));
  }
  // CAM_SVD_TRACE=1: time of the subspace rounds, the Rayleigh-Ritz
  // projection, the k x k core and the residual check of every check
  const char* tv = std::getenv("CAM_SVD_TRACE");
  const bool trace = tv && *tv && *tv != '0';
  auto tick = [&](double& acc, std::chrono::steady_clock::time_point& t) {
    if (!trace) return;
    HIP_CHECK(hipStreamSynchronize(s));
    const auto now = std::chrono::steady_clock::now();
    acc += std::chrono::duration<double>(now - t).count();
    t = now;
  };
  double t_round = 0, t_proj = 0, t_core = 0, t_res = 0;
  auto tp = std::chrono::steady_clock::now();
  // Between checks only the span of the iterates matters: one Cholesky QR
  // pass per multiplication keeps the condition of the next iterate near
  // that of A (its orthogonality error, about u cond^2, is irrelevant to
  // the span); where that pass breaks down the robust orthonormalisation
  // takes over.  The Rayleigh-Ritz step below always orthonormalises fully.
  G.ensure(static_cast<size_t>(k) * k);
  dinfo.ensure(1);
  DevBuf<double> dratio(1);
  auto span_step = [&](cplx* X, int64_t rows) {
    if (!cholqr_pass(ctx, X, rows, k, G.get(), dinfo.get(), 0.0, 1e-7, dratio.get()))
      orthonormalise(ctx, X, rows, k, G, dinfo, tau);
  };
  for (;;) {
    // before the first residual check: four subspace rounds on blocks of
    // more than 2^24 elements (they have needed at least six), else two, or
    // the rounds the last block of this shape needed; then two between checks
    const int nr = q == 0 ? first : 2;
    for (int r = 0; r < nr; ++r, ++q) {
      span_step(Y.get(), m);
      gemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, n, k, m, A, m, Y.get(), m, W.get(), n);
      span_step(W.get(), n);
      gemm(h, rocblas_operation_none, rocblas_operation_none, m, k, n, A, m, W.get(), n, Y.get(), m);
    }
    tick(t_round, tp);
    // Rayleigh-Ritz on span(Y)
    DevBuf<cplx> Q(m * k);
    HIP_CHECK(hipMemcpyAsync(Q.get(), Y.get(), m * k * sizeof(cplx), hipMemcpyDeviceToDevice, s));
    orthonormalise(ctx, Q.get(), m, k, G, dinfo, tau);
    // A^H Q = Q2 R: Q2 by Cholesky QR of a copy, R^H = (A^H Q)^H Q2
    gemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, n, k, m, A, m, Q.get(), m, W.get(), n);
    HIP_CHECK(hipMemcpyAsync(Wc.get(), W.get(), n * k * sizeof(cplx), hipMemcpyDeviceToDevice, s));
    orthonormalise(ctx, W.get(), n, k, G, dinfo, tau);
    gemm(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, k, k, n, Wc.get(), n, W.get(), n, R.get(),
         k);
    // k x k core R^H = Ur S Vr^H through the Hermitian eigenproblem of
    // [[0, R^H], [R, 0]] (divide and conquer; no squared condition)
    tick(t_proj, tp);
    std::vector<double> core;
    jordan_wielandt_svd(ctx, R.get(), k, k, k, k, core, Ur.get(), Vrh.get());
    dS.upload(core.data(), k, s);
    tick(t_core, tp);
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
    for (int64_t j = 0; j < naccept; ++j) worst = std::max(worst, std::sqrt(res[j]));
    tick(t_res, tp);
    if (trace)
      std::fprintf(stderr, "truncated_svd %lld x %lld k=%lld: after %d rounds worst residual of the first %lld "
                   "%.3e (tol %.3e); rounds %.3f s, projection %.3f s, core %.3f s, residual %.3f s\n",
                   static_cast<long long>(m), static_cast<long long>(n), static_cast<long long>(k), q,
                   static_cast<long long>(naccept), worst, tol, t_round, t_proj, t_core, t_res);
    if (rounds) *rounds = q;
    if (worst <= tol) {
      if (res_out) {
        res_out->resize(ncheck);
        for (int64_t j = 0; j < ncheck; ++j) (*res_out)[j] = std::sqrt(res[j]);
      }
      if (last_res) {
        // residual of the last computed triplet (the slowest to converge):
        // bounds how far its value may sit below a singular value of A
        *last_res = 0.0;
        if (ncheck < k) {
          gemm(h, rocblas_operation_none, rocblas_operation_conjugate_transpose, m, 1, n, A, m, Vh.get() + (k - 1),
               k, E.get(), m);
          hipLaunchKernelGGL(ritz_residual_kernel, dim3(1), dim3(256), 0, s, reinterpret_cast<double2*>(E.get()), m,
                             reinterpret_cast<const double2*>(U.get() + (k - 1) * m), dS.get() + (k - 1),
                             dres.get());
          HIP_CHECK(hipGetLastError());
          double r2 = 0.0;
          HIP_CHECK(hipMemcpyAsync(&r2, dres.get(), sizeof(double), hipMemcpyDeviceToHost, s));
          HIP_CHECK(hipStreamSynchronize(s));
          *last_res = std::sqrt(r2);
        }
      }
      // accepted at the first check: after three such blocks in a row the
      // hint is lowered by one round; otherwise it becomes the rounds needed
      auto& h = ctx.svd_round_hint[shape];
      if (q == first && q > first_default) {
        if (++h[1] >= 3) {
          h[0] = q - 1;
          h[1] = 0;
        } else {
          h[0] = q;
        }
      } else {
        h[0] = q;
        h[1] = 0;
      }
      return true;
    }
    if (q >= max_rounds) return false;
    // continue iterating from the current subspace
    HIP_CHECK(hipMemcpyAsync(Y.get(), U.get(), m * k * sizeof(cplx), hipMemcpyDeviceToDevice, s));
  }
}

}  // namespace cam
