// Device Lanczos for the lowest eigenpair of H_eff.
//
// Replaces arpack_eigs<TwoVertexComponents,SparseVectorWithRestriction>
// (arpack_interface.hpp:85-111, 174-237) as called from
// TwoVertexComponents::HED (DMRG_routines.cpp:835-896): nev=1, which="SR",
// Krylov dimension ncv (reference 20, arpack_interface.cpp:36-37),
// convergence when the Ritz residual |beta_m s_m| <= tol*|theta| (ARPACK's
// criterion), tol=0.1 for the non-converged growth steps
// (arpack_interface.cpp:28-31), machine epsilon when converging.  Restarts
// use the current Ritz vector (ARPACK restarts implicitly; the converged
// pair is the same).
//
// All vectors stay on the device in the block layout; Krylov vectors are
// kept fully orthogonal: each new vector is first projected off the last
// two basis vectors (the large alpha and beta components of the three-term
// recurrence), then off the whole basis in one classical Gram-Schmidt pass
// (FP64 sums); the tridiagonal eigenproblem is solved in FP64 on the device
// with rocSOLVER dsteqr.
//
// Mixed precision: each restart starts from the FP64 Ritz vector x, whose
// FP64 matvec gives the exact Rayleigh quotient and residual r; the Krylov
// space {x, r, H r, ...} is then extended with FP32 split-complex matvecs.
// FP32 errors enter only the correction, so the FP64 residual keeps
// decreasing; when it stalls, or once it is within 1e3*eps*|theta|, the
// remaining restarts run in FP64.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <limits>

#include "solvers.hpp"

namespace cam {

void* Ctx::blas_handle() {
  if (!blas) {
    rocblas_handle h;
    CAM_REQUIRE(rocblas_create_handle(&h) == rocblas_status_success, "rocblas_create_handle failed");
    CAM_REQUIRE(rocblas_set_stream(h, stream) == rocblas_status_success, "rocblas_set_stream failed");
    blas = h;
  }
  return blas;
}

Ctx::~Ctx() {
  if (blas) rocblas_destroy_handle(static_cast<rocblas_handle>(blas));
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {

// Lanczos step tail on the device: beta = |w| from the norm slot; the step
// breaks down when beta <= 16 eps max(1, |alpha|), alpha = Re of the summed
// projection coefficients on v_j, c1[j] (local pass) + c2[j] (full pass).  Writes v_{j+1} = w / beta (zero after a
// breakdown) and the decision {beta, flag} for the host.
template <typename T>
__global__ void lanczos_next_kernel(const T* __restrict__ w, T* __restrict__ next, int64_t n,
                                    const double2* __restrict__ nrm, const double2* __restrict__ c1,
                                    const double2* __restrict__ c2, double eps16, double2* __restrict__ dec) {
  const double alpha = c1->x + c2->x;
  const double beta = sqrt(fmax(0.0, nrm->x));
  const bool brk = beta <= eps16 * fmax(1.0, fabs(alpha));
  const double inv = brk ? 0.0 : 1.0 / beta;
  if (blockIdx.x == 0 && threadIdx.x == 0) *dec = make_double2(beta, brk ? 1.0 : 0.0);
  if (!next) return;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const T v = w[i];
    next[i] = T{static_cast<decltype(v.x)>(v.x * inv), static_cast<decltype(v.x)>(v.y * inv)};
  }
}

struct Lanczos {
  Heff& H;
  hipStream_t s;
  int64_t n;
  int m;                      // Krylov dimension
  DevView<cplx> V;            // (m+1) x n, in the context's Krylov workspace
  DevView<cplx> w;            // work (the Ritz vector lives in the caller's buffer)
  // FP32 cycles keep their Krylov vectors in FP32 (split complex), in the
  // first half of the same workspace: columns 0..m and the work vector w32,
  // all below the FP64 work vector w
  float2* V32 = nullptr;
  float2* w32 = nullptr;
  DevBuf<cplx> coef, partial;
  DevBuf<cplx> C1, C2, Nrm, Dec;   // per-step projection coefficients, norms, decisions
  DevBuf<double> dD, dE, dC;
  DevBuf<int> dinfo;
  hipGraphExec_t cycle_exec[2] = {nullptr, nullptr};
  bool use_graphs = true;
  int mv64 = 0, mv32 = 0;

  Lanczos(Heff& h, int ncv) : H(h), s(h.ctx->stream), n(h.n) {
    m = std::max(1, static_cast<int>(std::min<int64_t>(ncv, n)));
    CAM_REQUIRE(m + 1 <= 64, "ncv too large (max 63)");
    h.ctx->ws_krylov.ensure(static_cast<size_t>(n) * (m + 2));
    V.p = h.ctx->ws_krylov.get();
    w.p = V.p + static_cast<size_t>(n) * (m + 1);
    V32 = reinterpret_cast<float2*>(V.p);
    w32 = V32 + static_cast<size_t>(n) * (m + 1);
    coef.alloc(64);
    partial.alloc(static_cast<size_t>(multi_dot_blocks(n)) * 64);
    C1.alloc(64 * 64);
    C2.alloc(64 * 64);
    Nrm.alloc(64);
    Dec.alloc(64);
    dD.alloc(64);
    dE.alloc(64);
    dC.alloc(64 * 64);
    dinfo.alloc(1);
    const char* e = std::getenv("CAM_NO_GRAPH");
    use_graphs = !(e && *e && *e != '0');
  }
  ~Lanczos() {
    for (auto& g : cycle_exec)
      if (g) (void)hipGraphExecDestroy(g);
  }
  Lanczos(const Lanczos&) = delete;
  Lanczos& operator=(const Lanczos&) = delete;
  cplx* col(int j) { return V.get() + static_cast<int64_t>(j) * n; }
  float2* col32(int j) { return V32 + static_cast<int64_t>(j) * n; }

  // Steps j = 1 .. m-1 of one Krylov cycle with every scalar kept on the
  // device (col(1) must hold the normalised residual of step 0): matvec,
  // projection off v_{j-1}, v_j, one full Gram-Schmidt pass, norm,
  // normalisation.  No host synchronisation, so the whole sequence is one
  // hipGraph.
  //
  // H v_j lies in span(v_{j-1}, v_j, v_{j+1}) up to rounding (and, in an
  // FP32 cycle, the FP32 error of the matvec), so the local pass removes
  // everything large; what is left along the older vectors is small next to
  // |w| unless the space is nearly invariant (which the breakdown test
  // catches), so the single full pass suffers no cancellation and restores
  // orthogonality to working precision -- where a full pass applied to
  // H v_j directly cancels |alpha| / beta of the vector and needs a second
  // pass.  Per step this streams 2j + 15 + ceil((j+1)/8) vectors (the dot
  // sweeps re-read w once per eight basis vectors) instead of the
  // 4j + 12 + 2 ceil((j+1)/8) of two full passes.
  void enqueue_cycle(int prec) {
    const double eps16 = 16 * std::numeric_limits<double>::epsilon();
    if (prec == 1) {
      // FP32 vectors (col32(0) = x, col32(1) = the normalised residual):
      // FP32 matvecs and Gram-Schmidt on FP32 storage, FP64 sums
      for (int j = 1; j < m; ++j) {
        H.apply_blocked32(col32(j), w32);
        cplx* c1 = C1.get() + j * 64;
        cplx* c2 = C2.get() + j * 64;
        launch_multi_dot32(col32(j - 1), n, 2, w32, n, partial.get(), c1 + (j - 1), s);
        launch_multi_axpy32(col32(j - 1), n, 2, c1 + (j - 1), w32, n, s);
        launch_multi_dot32(V32, n, j + 1, w32, n, partial.get(), c2, s);
        launch_multi_axpy32(V32, n, j + 1, c2, w32, n, s);
        launch_multi_dot32(w32, n, 1, w32, n, partial.get(), Nrm.get() + j, s);
        const int grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(2048, (n + 255) / 256)));
        hipLaunchKernelGGL(lanczos_next_kernel<float2>, dim3(grid), dim3(256), 0, s, w32,
                           j + 1 < m ? col32(j + 1) : nullptr, n, reinterpret_cast<const double2*>(Nrm.get() + j),
                           reinterpret_cast<const double2*>(c1 + j), reinterpret_cast<const double2*>(c2 + j),
                           eps16, reinterpret_cast<double2*>(Dec.get() + j));
        HIP_CHECK(hipGetLastError());
      }
      return;
    }
    for (int j = 1; j < m; ++j) {
      H.apply(col(j), w.get(), prec);
      cplx* c1 = C1.get() + j * 64;
      cplx* c2 = C2.get() + j * 64;
      // local pass c = [v_{j-1} v_j]^H w, w -= [v_{j-1} v_j] c; full pass
      // c = V^H w, w -= V c; then |w|^2
      launch_multi_dot(col(j - 1), n, 2, w.get(), n, partial.get(), c1 + (j - 1), s);
      launch_multi_axpy(col(j - 1), n, 2, c1 + (j - 1), w.get(), n, s);
      launch_multi_dot(V.get(), n, j + 1, w.get(), n, partial.get(), c2, s);
      launch_multi_axpy(V.get(), n, j + 1, c2, w.get(), n, s);
      launch_multi_dot(w.get(), n, 1, w.get(), n, partial.get(), Nrm.get() + j, s);
      const int grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(2048, (n + 255) / 256)));
      hipLaunchKernelGGL(lanczos_next_kernel<double2>, dim3(grid), dim3(256), 0, s,
                         reinterpret_cast<const double2*>(w.get()),
                         j + 1 < m ? reinterpret_cast<double2*>(col(j + 1)) : nullptr, n,
                         reinterpret_cast<const double2*>(Nrm.get() + j),
                         reinterpret_cast<const double2*>(c1 + j), reinterpret_cast<const double2*>(c2 + j),
                         eps16, reinterpret_cast<double2*>(Dec.get() + j));
      HIP_CHECK(hipGetLastError());
    }
  }
  // runs the cycle (captured once per precision, then replayed) and returns
  // alpha_1.., beta_1.. up to the first breakdown
  bool run_cycle(int prec, std::vector<double>& alpha, std::vector<double>& beta, int& k) {
    if (m > 1) {
      // launch configurations are fixed before they are captured
      if (!H.tuned(prec)) {
        if (prec == 1) H.apply_blocked32(col32(1), w32);
        else H.apply(col(1), w.get(), prec);
      }
      hipGraphExec_t& ge = cycle_exec[prec];
      if (!ge && use_graphs) {
        HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
          enqueue_cycle(prec);
        } catch (...) {
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        HIP_CHECK(hipStreamEndCapture(s, &g));
        const hipError_t st = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIP_CHECK(st);
      }
      if (ge) HIP_CHECK(hipGraphLaunch(ge, s));
      else enqueue_cycle(prec);
    }
    std::vector<cplx> c1(64 * 64), c2(64 * 64), dec(64);
    HIP_CHECK(hipMemcpyAsync(c1.data(), C1.get(), c1.size() * sizeof(cplx), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(c2.data(), C2.get(), c2.size() * sizeof(cplx), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(dec.data(), Dec.get(), dec.size() * sizeof(cplx), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    bool invariant = false;
    for (int j = 1; j < m; ++j) {
      if (prec == 0) ++mv64; else ++mv32;
      alpha.push_back(c1[j * 64 + j].re + c2[j * 64 + j].re);
      k = j + 1;
      if (dec[j].im != 0.0) {
        invariant = true;
        break;
      }
      beta.push_back(dec[j].re);
    }
    return invariant;
  }

  void matvec(const cplx* in, cplx* out, int prec) {
    H.apply(in, out, prec);
    if (prec == 0) ++mv64; else ++mv32;
  }
  std::vector<std::complex<double>> dots(int k, const cplx* vec) {
    launch_multi_dot(V.get(), n, k, vec, n, partial.get(), coef.get(), s);
    std::vector<std::complex<double>> c(k);
    HIP_CHECK(hipMemcpyAsync(c.data(), coef.get(), k * sizeof(cplx), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return c;
  }
  double norm(const cplx* vec) {
    // <vec,vec> through the multi-dot with V := vec
    launch_multi_dot(vec, n, 1, vec, n, partial.get(), coef.get(), s);
    std::complex<double> c;
    HIP_CHECK(hipMemcpyAsync(&c, coef.get(), sizeof(cplx), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return std::sqrt(std::max(0.0, c.real()));
  }
  // w -= V[:, :k] (V^H w), twice; returns the first-pass coefficients + corrections
  std::vector<std::complex<double>> orthogonalize(int k, cplx* vec) {
    std::vector<std::complex<double>> tot(k, 0.0);
    for (int pass = 0; pass < 2; ++pass) {
      auto c = dots(k, vec);
      for (int j = 0; j < k; ++j) tot[j] += c[j];
      HIP_CHECK(hipMemcpyAsync(coef.get(), c.data(), k * sizeof(cplx), hipMemcpyHostToDevice, s));
      launch_multi_axpy(V.get(), n, k, coef.get(), vec, n, s);
    }
    return tot;
  }
  // lowest eigenpair of the symmetric tridiagonal (alpha, beta) on the device
  double tridiag_lowest(const std::vector<double>& a, const std::vector<double>& b, std::vector<double>& vec) {
    const int k = static_cast<int>(a.size());
    rocblas_handle h = static_cast<rocblas_handle>(H.ctx->blas_handle());
    std::vector<double> e(std::max(k - 1, 1), 0.0);
    for (int i = 0; i + 1 < k; ++i) e[i] = b[i];
    HIP_CHECK(hipMemcpyAsync(dD.get(), a.data(), k * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(dE.get(), e.data(), e.size() * sizeof(double), hipMemcpyHostToDevice, s));
    const rocblas_status st =
        rocsolver_dsteqr(h, rocblas_evect_tridiagonal, k, dD.get(), dE.get(), dC.get(), k, dinfo.get());
    CAM_REQUIRE(st == rocblas_status_success, "rocsolver_dsteqr failed");
    std::vector<double> evals(k), C(static_cast<size_t>(k) * k);
    int info = 0;
    HIP_CHECK(hipMemcpyAsync(evals.data(), dD.get(), k * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(C.data(), dC.get(), C.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(&info, dinfo.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    CAM_REQUIRE(info == 0, "tridiagonal eigensolver did not converge");
    int lo = 0;
    for (int i = 1; i < k; ++i)
      if (evals[i] < evals[lo]) lo = i;
    vec.assign(C.begin() + static_cast<int64_t>(lo) * k, C.begin() + static_cast<int64_t>(lo + 1) * k);
    return evals[lo];
  }
};

}  // namespace

namespace {

__global__ void absmax_partial_kernel(const double2* __restrict__ x, int64_t n, double* pv, int64_t* pi) {
  __shared__ double sv[256];
  __shared__ int64_t si[256];
  double best = -1.0;
  int64_t bi = 0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double2 v = x[i];
    const double a = v.x * v.x + v.y * v.y;
    if (a > best || (a == best && i < bi)) { best = a; bi = i; }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const double a = sv[threadIdx.x + o];
      const int64_t b = si[threadIdx.x + o];
      if (a > sv[threadIdx.x] || (a == sv[threadIdx.x] && b < si[threadIdx.x])) {
        sv[threadIdx.x] = a;
        si[threadIdx.x] = b;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { pv[blockIdx.x] = sv[0]; pi[blockIdx.x] = si[0]; }
}

// x <- FixComplexPrecision(x * phase) with |x*phase| <= thr set to zero
__global__ void gauge_kernel(double2* __restrict__ x, int64_t n, double pr, double pim, double thr,
                             double eps) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double2 v = x[i];
    double re = v.x * pr - v.y * pim, im = v.x * pim + v.y * pr;
    if (re * re + im * im > thr * thr) {
      if (re != 0.0 && fabs(im / re) < eps) im = 0.0;
      else if (im != 0.0 && fabs(re / im) < eps) re = 0.0;
    } else {
      re = 0.0;
      im = 0.0;
    }
    x[i] = make_double2(re, im);
  }
}

}  // namespace

// Phase gauge of DMRG_routines.cpp:870-890 on a device vector: the
// largest-modulus entry made real positive, entries below eps*max dropped,
// FixComplexPrecision (sparse_interface.cpp:28-32) on the rest.  In block
// layout the "largest entry" is the same element as in reference order;
// ties are broken by block-layout position.
void gauge_fix(cplx* x, int64_t n, hipStream_t s) {
  const int nb = 256;
  DevBuf<double> pv(nb);
  DevBuf<int64_t> pi(nb);
  hipLaunchKernelGGL(absmax_partial_kernel, dim3(nb), dim3(256), 0, s, reinterpret_cast<const double2*>(x), n,
                     pv.get(), pi.get());
  HIP_CHECK(hipGetLastError());
  std::vector<double> hv(nb);
  std::vector<int64_t> hi(nb);
  HIP_CHECK(hipMemcpyAsync(hv.data(), pv.get(), nb * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(hi.data(), pi.get(), nb * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  int b = 0;
  for (int k = 1; k < nb; ++k)
    if (hv[k] > hv[b] || (hv[k] == hv[b] && hi[k] < hi[b])) b = k;
  cplx mv;
  HIP_CHECK(hipMemcpy(&mv, x + hi[b], sizeof(cplx), hipMemcpyDeviceToHost));
  const double mx = std::sqrt(mv.re * mv.re + mv.im * mv.im);
  if (mx == 0.0) return;
  // phase = mx / mv
  const double pr = mv.re / mx, pim = -mv.im / mx;
  const double eps = std::numeric_limits<double>::epsilon();
  hipLaunchKernelGGL(gauge_kernel, dim3(1024), dim3(256), 0, s, reinterpret_cast<double2*>(x), n, pr, pim,
                     eps * mx, eps);
  HIP_CHECK(hipGetLastError());
}

double lanczos_lowest(Heff& H, cplx* xout, bool have_start, bool converge, double tol, int ncv,
                      bool mixed, double* info) {
  const int64_t n = H.n;
  CAM_REQUIRE(n > 0, "empty eigenproblem");
  hipStream_t s = H.ctx->stream;
  const double eps = std::numeric_limits<double>::epsilon();
  double tol_eff = converge ? (tol > 0 ? tol : eps) : (tol > 0 ? tol : 0.1);
  // the residual of an FP64 matvec cannot go far below ~eps*|H|; the
  // converged criterion is taken as max(tol, 64 eps)
  tol_eff = std::max(tol_eff, 64 * eps);
  ncv = ncv > 0 ? ncv : 20;
  if (mixed) H.prepare_fp32();   // one-time conversions stay out of the captured cycle
  Lanczos L(H, ncv);
  // start vector (SparseVectorWithRestriction -> DumbExtractWithZeros); a
  // zero guess takes the info=-9 path: random start
  if (have_start) {
    have_start = L.norm(xout) > 0.0;
  }
  if (!have_start) launch_random_fill(xout, n, 0x1234ull, s);
  {
    const double nx = L.norm(xout);
    launch_scale(cplx{1.0 / nx, 0.0}, xout, n, s);
  }
  // ARPACK's maxiter: 50 implicit restarts converged, 2 for growth steps
  // (arpack_interface.cpp:28); explicit restarts need more when converging
  const int max_restarts = converge ? 200 : 2;
  double theta = 0.0, resid = 0.0, prev_resid = 1e300;
  int restarts = 0;
  bool use32 = mixed;
  while (true) {
    const int m = L.m;
    launch_copy(xout, L.col(0), n, s);
    // step 0 in FP64: exact Rayleigh quotient and residual of x
    L.matvec(L.col(0), L.w.get(), 0);
    auto c0 = L.orthogonalize(1, L.w.get());
    theta = c0[0].real();
    resid = L.norm(L.w.get());
    // like ARPACK, convergence is judged only once a full Krylov space has
    // been built: a good start vector still gets one cycle of ncv vectors
    if (restarts >= 1 && (resid <= tol_eff * std::max(std::fabs(theta), 1e-300) || restarts >= max_restarts))
      break;
    if (use32 && restarts > 1 && (resid > 0.5 * prev_resid || resid < 1e3 * eps * std::fabs(theta)))
      use32 = false;   // final FP64 Lanczos steps
    prev_resid = resid;
    std::vector<double> alpha{theta}, beta;
    bool invariant = resid <= 16 * eps * std::max(1.0, std::fabs(theta));
    int k = 1;
    const bool cycle32 = use32;
    if (!invariant) {
      beta.push_back(resid);
      launch_scale(cplx{1.0 / resid, 0.0}, L.w.get(), n, s);
      if (cycle32) {
        // x and the residual in FP32 (col(0)'s FP64 copy of x is dead now)
        launch_to_fp32(xout, L.col32(0), n, s);
        launch_to_fp32(L.w.get(), L.col32(1), n, s);
      } else {
        launch_copy(L.w.get(), L.col(1), n, s);
      }
      invariant = L.run_cycle(cycle32 ? 1 : 0, alpha, beta, k);
    }
    std::vector<double> svec;
    beta.resize(k, 0.0);
    const double th = L.tridiag_lowest(alpha, beta, svec);
    const double ritz_res = invariant ? 0.0 : std::fabs(beta[k - 1] * svec[k - 1]);
    std::vector<std::complex<double>> cs(k);
    for (int j = 0; j < k; ++j) cs[j] = svec[j];
    HIP_CHECK(hipMemcpyAsync(L.coef.get(), cs.data(), k * sizeof(cplx), hipMemcpyHostToDevice, s));
    if (cycle32) {
      // x <- s_0 x + sum_{j>=1} s_j v_j: the FP64 x keeps its precision and
      // only the correction comes from the FP32 vectors
      launch_scale(cplx{svec[0], 0.0}, xout, n, s);
      if (k > 1) launch_accumulate32(L.col32(1), n, k - 1, L.coef.get() + 1, xout, n, s);
    } else {
      launch_combine(L.V.get(), n, k, L.coef.get(), xout, n, s);
    }
    const double nx = L.norm(xout);
    launch_scale(cplx{1.0 / nx, 0.0}, xout, n, s);
    ++restarts;
    // FP64 runs stop on ARPACK's Ritz estimate; mixed runs re-check the new
    // vector with an exact FP64 matvec at the top of the loop
    if (!use32 && (invariant || ritz_res <= tol_eff * std::max(std::fabs(th), 1e-300))) {
      theta = th;
      resid = ritz_res;
      break;
    }
  }
  if (info) {
    info[0] = L.mv64;
    info[1] = L.mv32;
    info[2] = restarts;
    info[3] = resid;
  }
  gauge_fix(xout, n, s);
  HIP_CHECK(hipStreamSynchronize(s));
  return theta;
}

namespace {

__global__ void unit_vector_kernel(double2* x, int64_t n, int64_t j) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    x[i] = make_double2(i == j ? 1.0 : 0.0, 0.0);
}

// TranslationBlock<DenseMatrix>::TranslateRows cut (sparse_interface.hpp:414-425):
// entries with |x| < SPARSETOL are not stored
__global__ void cut_small_kernel(const double2* __restrict__ x, double2* __restrict__ y, int64_t n, double tol) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double2 v = x[i];
    y[i] = (sqrt(v.x * v.x + v.y * v.y) >= tol) ? v : make_double2(0.0, 0.0);
  }
}

}  // namespace

// Dense branch of TwoVertexComponents::HED (DMRG_routines.cpp:850-858): the
// whole H_eff is formed (column j = H_eff e_j, in the block layout, which is
// a permutation of the reference's allowed-index order) and diagonalised
// by hermitian_eig; the lowest eigenvector is returned with entries
// below SPARSETOL cut and no phase gauge, as the reference's LAPACK path.
double dense_lowest(Heff& H, cplx* xout, double* info) {
  const int64_t n = H.n;
  CAM_REQUIRE(n > 0 && n <= kSparseThreshold, "dense eigensolver: size out of range");
  hipStream_t s = H.ctx->stream;
  DevBuf<cplx> A(static_cast<size_t>(n) * n), e(n);
  DevBuf<double> D(n);
  const int grid = static_cast<int>((n + 255) / 256);
  for (int64_t j = 0; j < n; ++j) {
    hipLaunchKernelGGL(unit_vector_kernel, dim3(grid), dim3(256), 0, s, reinterpret_cast<double2*>(e.get()), n, j);
    HIP_CHECK(hipGetLastError());
    H.apply(e.get(), A.get() + j * n, 0);
  }
  hermitian_eig(*H.ctx, A.get(), n, D.get());
  hipLaunchKernelGGL(cut_small_kernel, dim3(grid), dim3(256), 0, s, reinterpret_cast<const double2*>(A.get()),
                     reinterpret_cast<double2*>(xout), n, std::numeric_limits<double>::epsilon());
  HIP_CHECK(hipGetLastError());
  DevBuf<cplx> partial(static_cast<size_t>(multi_dot_blocks(n)) * 64), nrm(1);
  launch_multi_dot(xout, n, 1, xout, n, partial.get(), nrm.get(), s);
  double e0 = 0.0;
  cplx nx{0.0, 0.0};
  HIP_CHECK(hipMemcpyAsync(&e0, D.get(), sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(&nx, nrm.get(), sizeof(cplx), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  CAM_REQUIRE(std::fabs(nx.re - 1.0) < 1e-6, "dense eigensolver returned a vector that is not normalised");
  if (info) {
    info[0] = static_cast<double>(n);
    info[1] = 0.0;
    info[2] = 0.0;
    info[3] = 0.0;
  }
  return e0;
}

double eigs_lowest(Heff& H, cplx* x, bool have_start, bool converge, double tol, int ncv, bool mixed,
                   double* info) {
  CAM_REQUIRE(ncv <= 0 || (ncv >= 2 && ncv <= 63), "eigensolver: ncv must be 2..63 (or <= 0 for 20)");
  // the dense route only without projector terms (DMRG_routines.cpp:852)
  if (H.n <= kSparseThreshold && H.nproj == 0) return dense_lowest(H, x, info);
  return lanczos_lowest(H, x, have_start, converge, tol, ncv, mixed, info);
}

void heff_lowest(Heff& H, const cplx* guess_ref, bool converge, double tol, int ncv, bool mixed,
                 double* eval, cplx* evec_ref, double* info) {
  const int64_t n = H.n;
  hipStream_t s = H.ctx->stream;
  H.build_perm();
  DevBuf<cplx> x(n), y(n);
  if (guess_ref) {
    y.upload(guess_ref, n, s);
    launch_scatter_map(y.get(), x.get(), H.perm.get(), n, s);
  }
  *eval = eigs_lowest(H, x.get(), guess_ref != nullptr, converge, tol, ncv, mixed, info);
  if (evec_ref) {
    launch_gather_map(x.get(), y.get(), H.perm.get(), n, s);
    HIP_CHECK(hipMemcpyAsync(evec_ref, y.get(), n * sizeof(cplx), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
}

}  // namespace cam
