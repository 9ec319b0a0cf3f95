// Transfer-matrix fixed points, Orthogonalise and the energy per vertex of a
// two-vertex unit cell (measurement.cpp:180-246, 371-397, 405-629;
// MPX_matrix.cpp:829-913; UnitCell.hpp:42-48).
//
// The transfer maps run as grouped MFMA launches over the QN blocks of the
// cell tensors (two per vertex); the largest-magnitude eigensolve that
// replaces ARPACK's znaupd (which = "LM", nev = 1, ncv = 20) is a
// Krylov-Schur iteration whose Arnoldi vectors stay on the device; only the
// m x m projected matrix (m = 20) comes to the host for its Schur form.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <limits>
#include <map>

#include "ortho.hpp"
#include "schur.hpp"

namespace cam {

namespace {

constexpr double kImagTol = 100.0 * std::numeric_limits<double>::epsilon();   // measurement.cpp:12
constexpr int kNcv = 20;             // arpack_interface.cpp:39 with nev = 1
constexpr int kMaxRestarts = 50;     // arpack_interface.cpp:28

int grid_of(int64_t n) { return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(8192, (n + 255) / 256))); }

// x[off_s + i + i n_s] = 1 for every sector s (the identity on the boundary)
__global__ void diag_set_kernel(double2* __restrict__ x, const int64_t* __restrict__ off,
                                const int64_t* __restrict__ nn, int nsec) {
  const int s = blockIdx.x;
  if (s >= nsec) return;
  const int64_t n = nn[s];
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) x[off[s] + i + i * n] = make_double2(1.0, 0.0);
}

// out = sum over sectors of the diagonal of x (one workgroup)
__global__ void diag_sum_kernel(const double2* __restrict__ x, const int64_t* __restrict__ off,
                                const int64_t* __restrict__ nn, int nsec, double2* __restrict__ out) {
  __shared__ double sr[256], si[256];
  double r = 0.0, im = 0.0;
  for (int s = 0; s < nsec; ++s) {
    const int64_t n = nn[s];
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
      const double2 v = x[off[s] + i + i * n];
      r += v.x;
      im += v.y;
    }
  }
  sr[threadIdx.x] = r;
  si[threadIdx.x] = im;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sr[threadIdx.x] += sr[threadIdx.x + w];
      si[threadIdx.x] += si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = make_double2(sr[0], si[0]);
}

// M_s = p x_s + (p x_s)^H for every n_s x n_s block s (blockIdx.y)
__global__ void hermitian_complete_kernel(const double2* __restrict__ x, double2* __restrict__ M,
                                          const int64_t* __restrict__ off, const int64_t* __restrict__ nn,
                                          double pr, double pi) {
  const int s = blockIdx.y;
  const int64_t n = nn[s], base = off[s];
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < n * n;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = e % n, j = e / n;
    const double2 a = x[base + e], b = x[base + j + i * n];
    // p a + conj(p b)
    M[base + e] = make_double2(pr * a.x - pi * a.y + pr * b.x - pi * b.y,
                               pr * a.y + pi * a.x - (pr * b.y + pi * b.x));
  }
}

// In place: every n_s x n_s block s (blockIdx.y) becomes its Hermitian part
// (x + x^H) / 2; each off-diagonal pair is owned by one thread.
__global__ void hermitian_part_kernel(double2* __restrict__ x, const int64_t* __restrict__ off,
                                      const int64_t* __restrict__ nn) {
  const int s = blockIdx.y;
  const int64_t n = nn[s], base = off[s];
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < n * n;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = e % n, j = e / n;
    if (i > j) continue;
    if (i == j) {
      x[base + e].y = 0.0;
      continue;
    }
    const double2 a = x[base + e], b = x[base + j + i * n];
    const double2 h = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
    x[base + e] = h;
    x[base + j + i * n] = make_double2(h.x, -h.y);
  }
}

// c[i].im = 0 for i < k (real inner products on the Hermitian subspace)
__global__ void real_part_kernel(double2* __restrict__ c, int k) {
  const int i = threadIdx.x;
  if (i < k) c[i].y = 0.0;
}

// From the eigenvectors V (n x n, column-major) of one block and its kept
// columns c0 .. c0+k-1 with roots r_i = sqrt|w_i|:
//   X (k x n) = diag(r) V_kept^H            (when X != nullptr)
//   Z (n x k) = V_kept diag(r^zpow)          (when Z != nullptr)
__global__ void sqrtdr_factor_kernel(const double2* __restrict__ V, int64_t n, int64_t c0, int64_t k,
                                     const double* __restrict__ roots, double2* __restrict__ X,
                                     double2* __restrict__ Z, int zpow) {
  for (int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; e < n * k;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t j = e % n, i = e / n;
    const double2 v = V[j + (c0 + i) * n];
    const double r = roots[i];
    if (X) X[i + j * k] = make_double2(r * v.x, -r * v.y);
    if (Z) {
      const double f = zpow > 0 ? r : 1.0 / r;
      Z[j + i * n] = make_double2(f * v.x, f * v.y);
    }
  }
}

zcplx to_z(const cplx& c) { return {c.re, c.im}; }
cplx to_c(const zcplx& z) { return {z.real(), z.imag()}; }

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Square (inward q, outward q) tensor with one block per sector of q: the
// layout of a transfer-map vector on the boundary index q.
TensorPtr square_tensor(const Rules& rules, const QIndex& q) {
  auto t = std::make_shared<BSTensor>();
  t->rules = rules;
  QIndex a = q, b = q;
  a.inward = true;
  a.physical = false;
  b.inward = false;
  b.physical = false;
  t->idx = {a, b};
  t->numrows = 1;
  for (int s = 0; s < q.nsec(); ++s) {
    Block bl;
    bl.sec = {s, s};
    bl.dims = {q.sec_size[s], q.sec_size[s]};
    t->blocks.push_back(std::move(bl));
  }
  t->finalize_layout();
  return t;
}

struct SectorTable {
  DevBuf<int64_t> off, n;
  int nsec = 0;
  SectorTable(const TransferOp::Bond& b, hipStream_t s) {
    nsec = static_cast<int>(b.n.size());
    off.alloc(std::max(nsec, 1));
    n.alloc(std::max(nsec, 1));
    off.upload(b.off.data(), b.off.size(), s);
    n.upload(b.n.data(), b.n.size(), s);
  }
};

TransferOp::Bond bond_of(const QIndex& q) {
  TransferOp::Bond b;
  for (int s = 0; s < q.nsec(); ++s) {
    b.n.push_back(q.sec_size[s]);
    b.off.push_back(b.total);
    b.total += q.sec_size[s] * q.sec_size[s];
  }
  return b;
}

}  // namespace

TransferOp::TransferOp(Ctx& c, const std::vector<TensorPtr>& cell) : ctx(&c) {
  const int L = static_cast<int>(cell.size());
  CAM_REQUIRE(L >= 1 && L <= 8, "transfer matrix: the unit cell must hold 1 to 8 matrices");
  for (const TensorPtr& A : cell) {
    CAM_REQUIRE(A && A->nidx() == 3 && A->idx[0].physical && A->idx[1].inward && !A->idx[1].physical &&
                    !A->idx[2].inward && !A->idx[2].physical,
                "transfer matrix: cell matrices must have indices (physical, inward, outward)");
  }
  for (int i = 1; i < L; ++i)
    CAM_REQUIRE(cell[i]->idx[1].same_charges(cell[i - 1]->idx[2]),
                "transfer matrix: consecutive cell matrices do not share a bond index");
  CAM_REQUIRE(cell[L - 1]->idx[2].size == cell[0]->idx[1].size,
              "Index dimensions mismatch between ends of translation unit!");
  CAM_REQUIRE(cell[L - 1]->idx[2].same_charges(cell[0]->idx[1]),
              "transfer matrix: the two boundary indices of the cell carry different charges");
  boundary = &cell[0]->idx[1];
  hipStream_t s = ctx->stream;
  for (int i = 0; i < L; ++i) bonds.push_back(bond_of(cell[i]->idx[1]));
  bonds.push_back(bond_of(cell[L - 1]->idx[2]));
  int64_t wmax = 1, midmax = 1;
  for (int i = 0; i < L; ++i) {
    P.push_back(reorder(*ctx, *cell[i], false, {1, 0, 2}, 2));
    wmax = std::max(wmax, P.back()->total);
    if (i > 0) midmax = std::max(midmax, bonds[i].total);
  }
  wbuf.alloc(static_cast<size_t>(wmax));
  if (L > 1) {
    mid[0].alloc(static_cast<size_t>(midmax));
    mid[1].alloc(static_cast<size_t>(midmax));
  }
  // operand addressing: space 1 = the input vector, space 2 = the output
  // vector, space 0 = absolute (site tensors, scratch, intermediates)
  auto off = [](int64_t o) { return reinterpret_cast<const void*>(o); };
  for (int side = 0; side < 2; ++side) {
    for (int t = 0; t < L; ++t) {
      const int i = side == 0 ? t : L - 1 - t;
      const BSTensor& Pi = *P[i];
      const Bond& bin = side == 0 ? bonds[i] : bonds[i + 1];
      const Bond& bout = side == 0 ? bonds[i + 1] : bonds[i];
      const int inspace = t == 0 ? 1 : 0;
      const int outspace = t == L - 1 ? 2 : 0;
      const cplx* inbase = t == 0 ? nullptr : mid[(t - 1) & 1].get();
      cplx* outbase = t == L - 1 ? nullptr : mid[t & 1].get();
      auto inptr = [&](int64_t o) { return inspace ? off(o) : static_cast<const void*>(inbase + o); };
      auto outptr = [&](int64_t o) {
        return outspace ? reinterpret_cast<void*>(o) : static_cast<void*>(outbase + o);
      };
      GemmPlan& g1 = plan[side][0][i];
      GemmPlan& g2 = plan[side][1][i];
      // first product: per block of P (sa, ss, sb)
      for (const Block& B : Pi.blocks) {
        const int64_t na = B.dims[0], ds = B.dims[1], nb = B.dims[2];
        cplx* wblk = wbuf.get() + B.offset;
        const cplx* pblk = Pi.data.get() + B.offset;
        if (side == 0) {
          // W[a', (s, b)] = v_sa[a', a] P[a, (s, b)]
          const int pi = g1.add_problem(wblk, na, static_cast<int>(na), static_cast<int>(ds * nb), 1, 0, 0);
          g1.add_segment(pi, inptr(bin.off[B.sec[0]]), na, 0, pblk, na, 0, static_cast<int>(na),
                         seg_spaces(inspace, 0));
        } else {
          // T[(a, s), c'] = P[(a, s), c] v_sb[c', c]
          const int pi = g1.add_problem(wblk, na * ds, static_cast<int>(na * ds), static_cast<int>(nb), 1, 0, 0);
          g1.add_segment(pi, pblk, na * ds, 0, inptr(bin.off[B.sec[2]]), nb, 0, static_cast<int>(nb),
                         kSegBNContig | seg_spaces(0, inspace));
        }
      }
      // second product: per output sector
      const int outpos = side == 0 ? 2 : 0;
      std::vector<std::vector<const Block*>> by(bout.n.size());
      for (const Block& B : Pi.blocks) by[B.sec[outpos]].push_back(&B);
      for (size_t so = 0; so < bout.n.size(); ++so) {
        const int64_t no = bout.n[so];
        if (no == 0) continue;
        if (by[so].empty()) {
          empty[side][i].push_back({bout.off[so], no * no});
          continue;
        }
        const int pi = g2.add_problem(outptr(bout.off[so]), no, static_cast<int>(no), static_cast<int>(no), 1, 0, 0,
                                      outspace);
        for (const Block* B : by[so]) {
          const int64_t na = B->dims[0], ds = B->dims[1], nb = B->dims[2];
          const cplx* wblk = wbuf.get() + B->offset;
          const cplx* pblk = Pi.data.get() + B->offset;
          if (side == 0) {
            // out[b', b] += P[(a', s), b']^* W[(a', s), b]
            g2.add_segment(pi, pblk, na * ds, 0, wblk, na * ds, 0, static_cast<int>(na * ds),
                           kSegAKContig | kSegConjA);
          } else {
            // out[a', a] += P[a', (s, c')]^* T[a, (s, c')]
            g2.add_segment(pi, pblk, na, 0, wblk, na, 0, static_cast<int>(ds * nb), kSegConjA | kSegBNContig);
          }
        }
      }
      for (GemmPlan* g : {&g1, &g2}) {
        if (g->probs.empty()) continue;
        g->choose_tiles(1);
        g->build_tiles();
        flops[side] += g->flops;
      }
    }
  }
  HIP_CHECK(hipStreamSynchronize(s));
}

void TransferOp::apply(int side, const cplx* in, cplx* out) {
  const int L = static_cast<int>(P.size());
  hipStream_t s = ctx->stream;
  GemmBases b{};
  b.p[1] = in;
  b.p[2] = out;
  for (int t = 0; t < L; ++t) {
    const int i = side == 0 ? t : L - 1 - t;
    cplx* dst = t == L - 1 ? out : mid[t & 1].get();
    for (const auto& r : empty[side][i]) launch_zero(dst + r.first, r.second * static_cast<int64_t>(sizeof(cplx)), s);
    for (int g = 0; g < 2; ++g) {
      GemmPlan& pl = plan[side][g][i];
      if (pl.probs.empty()) continue;
      if (!tuned[side]) autotune(pl, 0, s, &b);
      launch_grouped_gemm(pl, 0, s, &b);
    }
  }
  tuned[side] = true;
}

LeadingEig leading_eigenpair(Ctx& ctx, const std::function<void(const cplx*, cplx*)>& mv, int64_t n, cplx* x,
                             int m, int max_restarts, double tol, const std::function<void(cplx*)>& project) {
  hipStream_t s = ctx.stream;
  CAM_REQUIRE(n >= 1, "eigensolver: empty vector space");
  m = static_cast<int>(std::min<int64_t>(m, n));
  CAM_REQUIRE(m >= 1 && m <= 63, "eigensolver: Krylov dimension out of range");
  const int keep = std::max(1, m / 2);
  DevBuf<cplx> V(static_cast<size_t>(m + 1) * n), w(n), tmp(static_cast<size_t>(keep + 1) * n), coef(2 * 64 + 1),
      partial(static_cast<size_t>(multi_dot_blocks(n)) * 64), zdev(static_cast<size_t>(m) * m);
  auto col = [&](int j) { return V.get() + static_cast<int64_t>(j) * n; };
  auto dot_self = [&](const cplx* v) {
    launch_multi_dot(v, n, 1, v, n, partial.get(), coef.get() + 128, s);
    cplx h;
    HIP_CHECK(hipMemcpyAsync(&h, coef.get() + 128, sizeof(cplx), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return h.re;
  };
  const double nx = dot_self(x);
  CAM_REQUIRE(nx > 0.0, "eigensolver: zero start vector");
  launch_copy(x, col(0), n, s);
  launch_scale(cplx{1.0 / std::sqrt(nx), 0.0}, col(0), n, s);

  std::vector<zcplx> G(static_cast<size_t>(m) * m, 0.0), T, Z;
  std::vector<cplx> hc(129);
  LeadingEig out;
  int k = 0;
  for (int r = 0;; ++r) {
    int mm = m;
    double beta = 0.0;
    for (int j = k; j < m; ++j) {
      mv(col(j), w.get());
      ++out.matvecs;
      cplx* c1 = coef.get();
      cplx* c2 = coef.get() + 64;
      if (project) project(w.get());
      launch_multi_dot(V.get(), n, j + 1, w.get(), n, partial.get(), c1, s);
      if (project) hipLaunchKernelGGL(real_part_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<double2*>(c1), j + 1);
      launch_multi_axpy(V.get(), n, j + 1, c1, w.get(), n, s);
      launch_multi_dot(V.get(), n, j + 1, w.get(), n, partial.get(), c2, s);
      if (project) hipLaunchKernelGGL(real_part_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<double2*>(c2), j + 1);
      launch_multi_axpy(V.get(), n, j + 1, c2, w.get(), n, s);
      launch_multi_dot(w.get(), n, 1, w.get(), n, partial.get(), coef.get() + 128, s);
      HIP_CHECK(hipMemcpyAsync(hc.data(), coef.get(), 129 * sizeof(cplx), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      double cn = 0.0;
      for (int i = 0; i <= j; ++i) {
        const zcplx h = to_z(hc[i]) + to_z(hc[64 + i]);
        G[i + static_cast<size_t>(j) * m] = h;
        cn += std::norm(h);
      }
      beta = std::sqrt(std::max(hc[128].re, 0.0));
      if (beta <= 64.0 * std::numeric_limits<double>::epsilon() * std::sqrt(cn)) {
        mm = j + 1;   // invariant subspace: the Ritz values of G are exact
        beta = 0.0;
        break;
      }
      if (j + 1 < m) G[j + 1 + static_cast<size_t>(j) * m] = beta;
      launch_copy(w.get(), col(j + 1), n, s);
      launch_scale(cplx{1.0 / beta, 0.0}, col(j + 1), n, s);
    }
    T.assign(static_cast<size_t>(mm) * mm, 0.0);
    for (int j = 0; j < mm; ++j)
      for (int i = 0; i < mm; ++i) T[i + static_cast<size_t>(j) * mm] = G[i + static_cast<size_t>(j) * m];
    CAM_REQUIRE(complex_schur(mm, T, Z), "eigensolver: Schur decomposition of the projected matrix failed");
    // one position past the kept ones is ordered too, so that a conjugate
    // pair cut by the restart is recognised
    schur_order_by_magnitude(mm, T, Z, std::min(keep + 1, mm));
    const zcplx theta = T[0];
    const double res = beta * std::abs(Z[mm - 1]);
    out.value = to_c(theta);
    out.residual = std::abs(theta) > 0.0 ? res / std::abs(theta) : res;
    out.restarts = r;
    const bool conv = res <= tol * std::abs(theta) || mm < m;
    if (conv || r >= max_restarts) {
      std::vector<cplx> z0(mm);
      for (int i = 0; i < mm; ++i) z0[i] = to_c(Z[i]);
      zdev.upload(z0.data(), z0.size(), s);
      launch_combine(V.get(), n, mm, zdev.get(), x, n, s);
      const double nn = dot_self(x);
      launch_scale(cplx{1.0 / std::sqrt(nn), 0.0}, x, n, s);
      HIP_CHECK(hipStreamSynchronize(s));
      out.converged = conv;
      return out;
    }
    // thick restart: V[:, :kk] = V[:, :m] Q, V[:, kk] = V[:, m], with Q an
    // orthonormal basis of the invariant subspace of the kk leading Schur
    // vectors; G becomes Q^H G Q with the coupling row beta e_m^T Q
    int kk = keep;
    bool explicit_restart = false;
    std::vector<zcplx> Q(static_cast<size_t>(mm) * (kk + 1), 0.0);
    if (!project) {
      for (int j = 0; j < kk; ++j)
        for (int i = 0; i < mm; ++i) Q[i + static_cast<size_t>(j) * mm] = Z[i + static_cast<size_t>(j) * mm];
    } else {
      // real basis: keep conjugate pairs together, then pivoted Gram-Schmidt
      // on the real and imaginary parts of the kept Schur vectors
      const zcplx a = T[(kk - 1) + static_cast<size_t>(kk - 1) * mm];
      if (kk < mm - 1 && std::fabs(a.imag()) > 1e-12 * std::abs(a)) {
        const zcplx b = T[kk + static_cast<size_t>(kk) * mm];
        if (std::abs(b - std::conj(a)) <= 1e-8 * std::abs(a)) ++kk;
      }
      std::vector<std::vector<double>> cols;
      for (int j = 0; j < kk; ++j) {
        std::vector<double> re(mm), im(mm);
        for (int i = 0; i < mm; ++i) {
          re[i] = Z[i + static_cast<size_t>(j) * mm].real();
          im[i] = Z[i + static_cast<size_t>(j) * mm].imag();
        }
        cols.push_back(std::move(re));
        cols.push_back(std::move(im));
      }
      for (int r = 0; r < kk; ++r) {
        size_t best = 0;
        double bn = -1.0;
        for (size_t c = 0; c < cols.size(); ++c) {
          double nn2 = 0.0;
          for (double v : cols[c]) nn2 += v * v;
          if (nn2 > bn) {
            bn = nn2;
            best = c;
          }
        }
        CAM_REQUIRE(bn > 0.0, "eigensolver: degenerate restart basis");
        std::vector<double> q = cols[best];
        cols.erase(cols.begin() + static_cast<long>(best));
        for (int pass = 0; pass < 2; ++pass)
          for (int t = 0; t < r; ++t) {
            double d = 0.0;
            for (int i = 0; i < mm; ++i) d += Q[i + static_cast<size_t>(t) * mm].real() * q[i];
            for (int i = 0; i < mm; ++i) q[i] -= d * Q[i + static_cast<size_t>(t) * mm].real();
          }
        double nq = 0.0;
        for (double v : q) nq += v * v;
        nq = std::sqrt(nq);
        CAM_REQUIRE(nq > 0.0, "eigensolver: degenerate restart basis");
        for (int i = 0; i < mm; ++i) Q[i + static_cast<size_t>(r) * mm] = q[i] / nq;
        for (auto& c : cols) {
          double d = 0.0;
          for (int i = 0; i < mm; ++i) d += Q[i + static_cast<size_t>(r) * mm].real() * c[i];
          for (int i = 0; i < mm; ++i) c[i] -= d * Q[i + static_cast<size_t>(r) * mm].real();
        }
      }
    }
    // Gk = Q^H G Q (G still holds the projected matrix of this cycle)
    std::vector<zcplx> GQ(static_cast<size_t>(mm) * kk, 0.0), Gk(static_cast<size_t>(kk) * kk, 0.0);
    for (int j = 0; j < kk; ++j)
      for (int l = 0; l < mm; ++l) {
        const zcplx q = Q[l + static_cast<size_t>(j) * mm];
        if (q == 0.0) continue;
        for (int i = 0; i < mm; ++i) GQ[i + static_cast<size_t>(j) * mm] += G[i + static_cast<size_t>(l) * m] * q;
      }
    for (int j = 0; j < kk; ++j)
      for (int i = 0; i < kk; ++i) {
        zcplx acc = 0.0;
        for (int l = 0; l < mm; ++l)
          acc += std::conj(Q[l + static_cast<size_t>(i) * mm]) * GQ[l + static_cast<size_t>(j) * mm];
        Gk[i + static_cast<size_t>(j) * kk] = acc;
      }
    if (project) {
      // the real basis must span an invariant subspace (G Q = Q Gk); when
      // it does not (a conjugate pair the ordering could not keep together),
      // restart from the real dominant Ritz vector alone
      double rn = 0.0, gn = 0.0;
      for (const zcplx& g : G) gn += std::norm(g);
      for (int j = 0; j < kk; ++j)
        for (int i = 0; i < mm; ++i) {
          zcplx qg = 0.0;
          for (int l = 0; l < kk; ++l) qg += Q[i + static_cast<size_t>(l) * mm] * Gk[l + static_cast<size_t>(j) * kk];
          rn += std::norm(GQ[i + static_cast<size_t>(j) * mm] - qg);
        }
      if (rn > 1e-20 * gn) {
        explicit_restart = true;
        kk = 1;
        size_t imax = 0;
        for (int i = 1; i < mm; ++i)
          if (std::abs(Z[i]) > std::abs(Z[imax])) imax = static_cast<size_t>(i);
        const zcplx ph = std::abs(Z[imax]) > 0.0 ? std::conj(Z[imax]) / std::abs(Z[imax]) : zcplx(1.0);
        double nq = 0.0;
        for (int i = 0; i < mm; ++i) nq += std::norm((ph * Z[i]).real());
        nq = std::sqrt(nq);
        CAM_REQUIRE(nq > 0.0, "eigensolver: degenerate restart vector");
        for (int i = 0; i < mm; ++i) Q[i] = (ph * Z[i]).real() / nq;
        zcplx acc = 0.0;
        for (int l = 0; l < mm; ++l) {
          zcplx gq = 0.0;
          for (int i = 0; i < mm; ++i) gq += G[l + static_cast<size_t>(i) * m] * Q[i];
          acc += std::conj(Q[l]) * gq;
        }
        Gk.assign(1, acc);
      }
    }
    std::vector<cplx> zk(static_cast<size_t>(mm) * kk);
    for (int j = 0; j < kk; ++j)
      for (int i = 0; i < mm; ++i) zk[i + static_cast<size_t>(j) * mm] = to_c(Q[i + static_cast<size_t>(j) * mm]);
    zdev.upload(zk.data(), zk.size(), s);
    for (int j = 0; j < kk; ++j)
      launch_combine(V.get(), n, mm, zdev.get() + static_cast<int64_t>(j) * mm, tmp.get() + static_cast<int64_t>(j) * n,
                     n, s);
    launch_copy(tmp.get(), V.get(), static_cast<int64_t>(kk) * n, s);
    std::fill(G.begin(), G.end(), 0.0);
    if (explicit_restart) {
      k = 0;   // a fresh Arnoldi cycle from the (unit) dominant Ritz vector
      continue;
    }
    launch_copy(col(m), col(kk), n, s);
    for (int j = 0; j < kk; ++j) {
      for (int i = 0; i < kk; ++i) G[i + static_cast<size_t>(j) * m] = Gk[i + static_cast<size_t>(j) * kk];
      G[kk + static_cast<size_t>(j) * m] = beta * Q[mm - 1 + static_cast<size_t>(j) * mm];
    }
    if (project)
      for (auto& g : G) g = g.real();
    k = kk;
  }
}

void transfer_apply(Ctx& ctx, const std::vector<TensorPtr>& cell, int side, const cplx* in, cplx* out) {
  CAM_REQUIRE(side == 0 || side == 1, "transfer matrix: side must be 0 (left) or 1 (right)");
  hipStream_t s = ctx.stream;
  TransferOp T(ctx, cell);
  const QIndex& q = *T.boundary;
  const TransferOp::Bond& b = side == 0 ? T.bonds.front() : T.bonds.back();
  // allowed (r, c) in the reference's loop order (c outer, r inner) -> block layout
  std::vector<int64_t> perm;
  perm.reserve(static_cast<size_t>(b.total));
  for (int64_t c = 0; c < q.size; ++c)
    for (int64_t r = 0; r < q.size; ++r) {
      if (q.charges[r] != q.charges[c]) continue;
      const int sct = q.pos_sec[c];
      perm.push_back(b.off[sct] + q.pos_off[r] + b.n[sct] * q.pos_off[c]);
    }
  const int64_t n = b.total;
  CAM_REQUIRE(static_cast<int64_t>(perm.size()) == n, "transfer matrix: allowed entries do not fill the blocks");
  DevBuf<int64_t> dperm(std::max<int64_t>(n, 1));
  dperm.upload(perm.data(), perm.size(), s);
  DevBuf<cplx> hin(n), x(n), y(n);
  hin.upload(in, n, s);
  launch_scatter_map(hin.get(), x.get(), dperm.get(), n, s);
  T.apply(side, x.get(), y.get());
  launch_gather_map(y.get(), hin.get(), dperm.get(), n, s);
  HIP_CHECK(hipMemcpyAsync(out, hin.get(), n * sizeof(cplx), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

LeadingEig transfer_leading(Ctx& ctx, const std::vector<TensorPtr>& cell, int side, double tol, TensorPtr* M,
                            int ncv) {
  CAM_REQUIRE(side == 0 || side == 1, "transfer matrix: side must be 0 (left) or 1 (right)");
  hipStream_t s = ctx.stream;
  TransferOp T(ctx, cell);
  const TransferOp::Bond& b = side == 0 ? T.bonds.front() : T.bonds.back();
  const int64_t n = b.total;
  SectorTable tab(b, s);
  DevBuf<cplx> x(n);
  launch_zero(x.get(), n * static_cast<int64_t>(sizeof(cplx)), s);
  hipLaunchKernelGGL(diag_set_kernel, dim3(std::max(tab.nsec, 1)), dim3(256), 0, s,
                     reinterpret_cast<double2*>(x.get()), tab.off.get(), tab.n.get(), tab.nsec);
  HIP_CHECK(hipGetLastError());
  int64_t nmax = 1;
  for (int64_t v : b.n) nmax = std::max(nmax, v * v);
  // The maps send Hermitian matrices to Hermitian matrices and the fixed
  // point is Hermitian: the iteration runs on that real subspace (Hermitian
  // part of every product, real inner products), so the projected matrix is
  // real and a simple dominant eigenvalue comes out real to rounding.
  auto project = [&](cplx* v) {
    hipLaunchKernelGGL(hermitian_part_kernel, dim3(grid_of(nmax), std::max(tab.nsec, 1)), dim3(256), 0, s,
                       reinterpret_cast<double2*>(v), tab.off.get(), tab.n.get());
  };
  LeadingEig e = leading_eigenpair(
      ctx, [&](const cplx* in, cplx* out) { T.apply(side, in, out); }, n, x.get(), ncv > 0 ? ncv : kNcv, kMaxRestarts,
      tol, project);
  if (M) {
    DevBuf<cplx> tr(1);
    hipLaunchKernelGGL(diag_sum_kernel, dim3(1), dim3(256), 0, s, reinterpret_cast<const double2*>(x.get()),
                       tab.off.get(), tab.n.get(), tab.nsec, reinterpret_cast<double2*>(tr.get()));
    HIP_CHECK(hipGetLastError());
    cplx h;
    HIP_CHECK(hipMemcpyAsync(&h, tr.get(), sizeof(cplx), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    const double a = std::hypot(h.re, h.im);
    const double pr = a > 0.0 ? h.re / a : 1.0, pi = a > 0.0 ? -h.im / a : 0.0;   // |tr| / tr
    const QIndex& q = side == 0 ? cell.front()->idx[1] : cell.back()->idx[2];
    *M = square_tensor(cell.front()->rules, q);
    CAM_REQUIRE((*M)->total == n, "transfer matrix: fixed-point layout mismatch");
    hipLaunchKernelGGL(hermitian_complete_kernel, dim3(grid_of(nmax), std::max(tab.nsec, 1)), dim3(256), 0, s,
                       reinterpret_cast<const double2*>(x.get()), reinterpret_cast<double2*>((*M)->data.get()),
                       tab.off.get(), tab.n.get(), pr, pi);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
  }
  return e;
}

namespace {

// SqrtDR (MPX_matrix.cpp:829-913) of a Hermitian square tensor M (inward q,
// outward q): block eigensolves, eigenvalues with |w| <= 100 eps max|w|
// dropped with their vectors, mixed signs rejected.  Builds
//   X    = diag(sqrt|w|) V^H   indices (inward new, outward q)   [left]
//   Xinv = V diag(1/sqrt|w|)   indices (inward q, outward new)   [left]
//   Y    = V diag(sqrt|w|)     indices (inward q, outward new)   [right]
// The new index lists the kept vectors block by block, each carrying its
// block's charge (DeduceFromColumns on a block-diagonal matrix).
int64_t sqrt_dr(Ctx& ctx, const BSTensor& M, TensorPtr* X, TensorPtr* Xinv, TensorPtr* Y) {
  hipStream_t s = ctx.stream;
  const QIndex& q = M.idx[0];
  const int ns = q.nsec();
  std::vector<DevBuf<cplx>> vecs(ns);
  std::vector<std::vector<double>> vals(ns);
  double big = 0.0;
  for (int sct = 0; sct < ns; ++sct) {
    const int64_t n = q.sec_size[sct];
    if (!n) continue;
    const int b = M.find_block({sct, sct});
    CAM_REQUIRE(b >= 0, "SqrtDR: missing diagonal block");
    vecs[sct].alloc(static_cast<size_t>(n * n));
    launch_copy(M.data.get() + M.blocks[b].offset, vecs[sct].get(), n * n, s);
    DevBuf<double> w(n);
    hermitian_eig(ctx, vecs[sct].get(), n, w.get());
    vals[sct].resize(n);
    HIP_CHECK(hipMemcpyAsync(vals[sct].data(), w.get(), n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (double v : vals[sct]) big = std::max(big, std::fabs(v));
  }
  const double tol = big * 100.0 * std::numeric_limits<double>::epsilon();
  std::vector<int64_t> c0(ns, 0), kk(ns, 0);
  int64_t finite = 0, negative = 0;
  for (int sct = 0; sct < ns; ++sct) {
    int64_t first = -1, last = -1;
    for (int64_t j = 0; j < static_cast<int64_t>(vals[sct].size()); ++j)
      if (std::fabs(vals[sct][j]) > tol) {
        if (first < 0) first = j;
        last = j;
        ++finite;
        if (vals[sct][j] < 0.0) ++negative;
      }
    if (first >= 0) {
      c0[sct] = first;
      kk[sct] = last - first + 1;
    }
  }
  if (negative != 0 && negative != finite) CAM_THROW("SqrtDR: Negative eigenvalue, can't decompose into Rdagger D R!");
  for (int sct = 0; sct < ns; ++sct)
    for (int64_t j = c0[sct]; j < c0[sct] + kk[sct]; ++j)
      CAM_REQUIRE(std::fabs(vals[sct][j]) > tol, "SqrtDR: kept eigenvalues are not contiguous");
  QIndex nq;
  nq.inward = true;
  for (int sct = 0; sct < ns; ++sct)
    for (int64_t i = 0; i < kk[sct]; ++i) nq.charges.push_back(q.sec_q[sct]);
  nq.size = static_cast<int64_t>(nq.charges.size());
  CAM_REQUIRE(nq.size > 0, "SqrtDR: no finite eigenvalue");
  nq.build_sectors();
  QIndex qin = q, qout = q, nout = nq;
  qin.inward = true;
  qout.inward = false;
  nout.inward = false;
  auto make = [&](std::vector<QIndex> idx, bool new_first) {
    auto t = std::make_shared<BSTensor>();
    t->rules = M.rules;
    t->idx = std::move(idx);
    t->numrows = 1;
    for (int sct = 0; sct < ns; ++sct) {
      if (!kk[sct]) continue;
      const int sn = nq.find_sector(q.sec_q[sct]);
      Block bl;
      bl.sec = new_first ? std::vector<int32_t>{sn, sct} : std::vector<int32_t>{sct, sn};
      bl.dims = new_first ? std::vector<int64_t>{kk[sct], q.sec_size[sct]} : std::vector<int64_t>{q.sec_size[sct], kk[sct]};
      t->blocks.push_back(std::move(bl));
    }
    std::sort(t->blocks.begin(), t->blocks.end(),
              [&](const Block& a, const Block& b) { return t->code(a.sec) < t->code(b.sec); });
    t->finalize_layout();
    return t;
  };
  if (X) *X = make({nq, qout}, true);
  if (Xinv) *Xinv = make({qin, nout}, false);
  if (Y) *Y = make({qin, nout}, false);
  for (int sct = 0; sct < ns; ++sct) {
    if (!kk[sct]) continue;
    const int sn = nq.find_sector(q.sec_q[sct]);
    const int64_t n = q.sec_size[sct], k = kk[sct];
    std::vector<double> roots(k);
    for (int64_t i = 0; i < k; ++i) roots[i] = std::sqrt(std::fabs(vals[sct][c0[sct] + i]));
    DevBuf<double> droots(k);
    droots.upload(roots.data(), roots.size(), s);
    auto blk = [&](const TensorPtr& t, bool new_first) {
      const int b = t->find_block(new_first ? std::vector<int32_t>{sn, sct} : std::vector<int32_t>{sct, sn});
      return reinterpret_cast<double2*>(t->data.get() + t->blocks[b].offset);
    };
    const int grid = grid_of(n * k);
    if (X || Xinv)
      hipLaunchKernelGGL(sqrtdr_factor_kernel, dim3(grid), dim3(256), 0, s,
                         reinterpret_cast<const double2*>(vecs[sct].get()), n, c0[sct], k, droots.get(),
                         X ? blk(*X, true) : nullptr, Xinv ? blk(*Xinv, false) : nullptr, -1);
    if (Y)
      hipLaunchKernelGGL(sqrtdr_factor_kernel, dim3(grid), dim3(256), 0, s,
                         reinterpret_cast<const double2*>(vecs[sct].get()), n, c0[sct], k, droots.get(), nullptr,
                         blk(*Y, false), 1);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
  }
  return nq.size;
}

TensorPtr diagonal_tensor(Ctx& ctx, const QIndex& q, const std::vector<double>& values, const Rules& rules) {
  const int64_t n = q.size;
  CAM_REQUIRE(static_cast<int64_t>(values.size()) == n, "diagonal: size mismatch");
  std::vector<int64_t> p(n + 1), i(n);
  std::vector<double> x(2 * n, 0.0);
  for (int64_t e = 0; e < n; ++e) {
    p[e] = e;
    i[e] = e;
    x[2 * e] = values[e];
  }
  p[n] = n;
  QIndex a = q, b = q;
  a.inward = true;
  a.physical = false;
  b.inward = false;
  b.physical = false;
  return tensor_from_csc(rules, {a, b}, 1, n, n, p.data(), i.data(), x.data(), ctx.stream);
}

cplx scalar_of(Ctx& ctx, const BSTensor& t) {
  for (const QIndex& q : t.idx) CAM_REQUIRE(q.size == 1, "measurement did not close to a number");
  if (t.total == 0) return cplx{0.0, 0.0};
  cplx h;
  HIP_CHECK(hipMemcpyAsync(&h, t.data.get(), sizeof(cplx), hipMemcpyDeviceToHost, ctx.stream));
  HIP_CHECK(hipStreamSynchronize(ctx.stream));
  return h;
}

}  // namespace

namespace {

// Charge-paired blocks of (bra index) x (ket index): one block per charge
// present in both, rows bra, the layout of a mixed transfer-map vector.
struct PairBond {
  std::vector<int> bs, ks;           // bra / ket sector of each pair
  std::vector<int64_t> nb, nk, off;
  int64_t total = 0;
  std::map<Charge, int> by_charge;
};

PairBond pair_bond(const QIndex& bra, const QIndex& ket) {
  PairBond p;
  for (int sb = 0; sb < bra.nsec(); ++sb) {
    const int sk = ket.find_sector(bra.sec_q[sb]);
    if (sk < 0) continue;
    p.by_charge[bra.sec_q[sb]] = static_cast<int>(p.bs.size());
    p.bs.push_back(sb);
    p.ks.push_back(sk);
    p.nb.push_back(bra.sec_size[sb]);
    p.nk.push_back(ket.sec_size[sk]);
    p.off.push_back(p.total);
    p.total += bra.sec_size[sb] * ket.sec_size[sk];
  }
  return p;
}

// LeftComponentsMultiply with Bra != Ket (measurement.cpp:582-602): per
// vertex W = v . A (ket) and out = B^H . W (bra), two grouped MFMA launches.
struct MixedTransfer {
  Ctx* ctx;
  std::vector<TensorPtr> Pk, Pb;
  std::vector<PairBond> bonds;
  GemmPlan plan[2][8];
  std::vector<std::pair<int64_t, int64_t>> empty[8];
  DevBuf<cplx> wbuf, mid[2];

  MixedTransfer(Ctx& c, const std::vector<TensorPtr>& bra, const std::vector<TensorPtr>& ket) : ctx(&c) {
    const int L = static_cast<int>(ket.size());
    CAM_REQUIRE(L >= 1 && L <= 8 && bra.size() == ket.size(), "Bra and Ket fragments have different sizes!");
    for (int i = 0; i < L; ++i)
      for (const TensorPtr& A : {bra[i], ket[i]})
        CAM_REQUIRE(A && A->nidx() == 3 && A->idx[0].physical && A->idx[1].inward && !A->idx[2].inward,
                    "transfer matrix: cell matrices must have indices (physical, inward, outward)");
    CAM_REQUIRE(bra[L - 1]->idx[2].same_charges(bra[0]->idx[1]) && ket[L - 1]->idx[2].same_charges(ket[0]->idx[1]),
                "transfer matrix: the two boundary indices of a cell carry different charges");
    for (int i = 0; i < L; ++i) bonds.push_back(pair_bond(bra[i]->idx[1], ket[i]->idx[1]));
    bonds.push_back(pair_bond(bra[L - 1]->idx[2], ket[L - 1]->idx[2]));
    int64_t wtot = 1, midmax = 1;
    std::vector<std::vector<int64_t>> woff(L);
    for (int i = 0; i < L; ++i) {
      Pk.push_back(reorder(c, *ket[i], false, {1, 0, 2}, 2));
      Pb.push_back(reorder(c, *bra[i], false, {1, 0, 2}, 2));
      int64_t w = 0;
      const PairBond& bin = bonds[i];
      for (const Block& B : Pk[i]->blocks) {
        woff[i].push_back(w);
        const Charge& qa = Pk[i]->idx[0].sec_q[B.sec[0]];
        auto it = bin.by_charge.find(qa);
        if (it == bin.by_charge.end()) continue;
        w += bin.nb[it->second] * B.dims[1] * B.dims[2];
      }
      wtot = std::max(wtot, w);
      if (i > 0) midmax = std::max(midmax, bonds[i].total);
    }
    wbuf.alloc(static_cast<size_t>(wtot));
    if (L > 1) {
      mid[0].alloc(static_cast<size_t>(midmax));
      mid[1].alloc(static_cast<size_t>(midmax));
    }
    auto off = [](int64_t o) { return reinterpret_cast<const void*>(o); };
    for (int i = 0; i < L; ++i) {
      const PairBond& bin = bonds[i];
      const PairBond& bout = bonds[i + 1];
      const int inspace = i == 0 ? 1 : 0, outspace = i == L - 1 ? 2 : 0;
      const cplx* inbase = i == 0 ? nullptr : mid[(i - 1) & 1].get();
      cplx* outbase = i == L - 1 ? nullptr : mid[i & 1].get();
      auto inptr = [&](int64_t o) { return inspace ? off(o) : static_cast<const void*>(inbase + o); };
      auto outptr = [&](int64_t o) { return outspace ? reinterpret_cast<void*>(o) : static_cast<void*>(outbase + o); };
      const BSTensor& K = *Pk[i];
      const BSTensor& Bt = *Pb[i];
      std::vector<std::vector<std::pair<int, int>>> by(bout.bs.size());   // (ket block, bra block) per output pair
      for (size_t kb = 0; kb < K.blocks.size(); ++kb) {
        const Block& B = K.blocks[kb];
        const Charge& qa = K.idx[0].sec_q[B.sec[0]];
        const Charge& qb = K.idx[2].sec_q[B.sec[2]];
        auto ia = bin.by_charge.find(qa);
        if (ia == bin.by_charge.end()) continue;
        const int pa = ia->second;
        const int64_t nbr = bin.nb[pa], nka = B.dims[0], ds = B.dims[1], nkb = B.dims[2];
        // W[r, (s, b)] = v[r, a] A[a, (s, b)]
        const int pi = plan[0][i].add_problem(wbuf.get() + woff[i][kb], nbr, static_cast<int>(nbr),
                                              static_cast<int>(ds * nkb), 1, 0, 0);
        plan[0][i].add_segment(pi, inptr(bin.off[pa]), nbr, 0, K.data.get() + B.offset, nka, 0,
                               static_cast<int>(nka), seg_spaces(inspace, 0));
        auto ib = bout.by_charge.find(qb);
        if (ib == bout.by_charge.end()) continue;
        const int bb = Bt.find_block({bin.bs[pa], B.sec[1], bout.bs[ib->second]});
        if (bb < 0) continue;
        by[ib->second].push_back({static_cast<int>(kb), bb});
      }
      for (size_t po = 0; po < bout.bs.size(); ++po) {
        const int64_t mo = bout.nb[po], no = bout.nk[po];
        if (!mo || !no) continue;
        if (by[po].empty()) {
          empty[i].push_back({bout.off[po], mo * no});
          continue;
        }
        const int pi = plan[1][i].add_problem(outptr(bout.off[po]), mo, static_cast<int>(mo), static_cast<int>(no), 1,
                                              0, 0, outspace);
        for (const auto& kb_bb : by[po]) {
          const Block& B = K.blocks[kb_bb.first];
          const Block& Bb = Bt.blocks[kb_bb.second];
          const int64_t kk = Bb.dims[0] * Bb.dims[1];   // (r, s) of the bra block
          // out[r', b] += conj(Pb[(r, s), r']) W[(r, s), b]
          plan[1][i].add_segment(pi, Bt.data.get() + Bb.offset, kk, 0, wbuf.get() + woff[i][kb_bb.first], kk, 0,
                                 static_cast<int>(kk), kSegAKContig | kSegConjA);
          (void)B;
        }
      }
      for (int g = 0; g < 2; ++g)
        if (!plan[g][i].probs.empty()) {
          plan[g][i].choose_tiles(1);
          plan[g][i].build_tiles();
        }
    }
    HIP_CHECK(hipStreamSynchronize(c.stream));
  }

  void apply(const cplx* in, cplx* out) {
    const int L = static_cast<int>(Pk.size());
    hipStream_t s = ctx->stream;
    GemmBases b{};
    b.p[1] = in;
    b.p[2] = out;
    for (int i = 0; i < L; ++i) {
      cplx* dst = i == L - 1 ? out : mid[i & 1].get();
      for (const auto& r : empty[i]) launch_zero(dst + r.first, r.second * static_cast<int64_t>(sizeof(cplx)), s);
      for (int g = 0; g < 2; ++g)
        if (!plan[g][i].probs.empty()) launch_grouped_gemm(plan[g][i], 0, s, &b);
    }
  }
};

}  // namespace

cplx overlap(Ctx& ctx, const std::vector<TensorPtr>& bra, const std::vector<TensorPtr>& ket) {
  MixedTransfer T(ctx, bra, ket);
  const int64_t n = T.bonds.front().total;
  CAM_REQUIRE(n > 0, "Overlap: no common charge sector");
  CAM_REQUIRE(T.bonds.back().total == n, "Index dimensions mismatch between ends of translation unit!");
  DevBuf<cplx> x(n);
  launch_random_fill(x.get(), n, 0x5eedull, ctx.stream);
  const LeadingEig e = leading_eigenpair(
      ctx, [&](const cplx* in, cplx* out) { T.apply(in, out); }, n, x.get(), kNcv, kMaxRestarts, 1e-14);
  CAM_REQUIRE(e.converged || e.residual <= 1e-10, "Overlap: the transfer-matrix eigensolve did not converge");
  return e.value;
}

TensorPtr unit_cell_second(Ctx& ctx, const BSTensor& Vd, const std::vector<double>& values,
                           const std::vector<double>& previous) {
  CAM_REQUIRE(Vd.nidx() == 3, "UnitCell: the row matrix must have three indices");
  TensorPtr A1 = reorder(ctx, Vd, false, {1, 0, 2}, 2);
  scale_along_index(ctx, *A1, 1, values);
  std::vector<double> inv(previous.size());
  for (size_t k = 0; k < previous.size(); ++k) inv[k] = 1.0 / previous[k];
  scale_along_index(ctx, *A1, 2, inv);
  return A1;
}

namespace {

void check_leading(const LeadingEig& e, const char* side) {
  if (!e.converged && !(e.residual <= 1e-10))
    CAM_THROW(std::string("Orthogonalise: the ") + side + " transfer-matrix eigensolve did not converge");
  if (std::fabs(e.value.im) >= kImagTol * std::fabs(e.value.re))
    CAM_THROW(std::string("Orthogonalise: the leading ") + side +
              " eigenvalue has a non negligible imaginary part");
}

// The common end of Orthogonalise and OrthogonaliseInversionSymmetric
// (measurement.cpp:166-177, 229-245): SVD of XLY, values rescaled to unit
// square sum, A0' = U^H X A0 / sqrt(l), A1' = A1 X^-1 U.
void finish(Ctx& ctx, const TensorPtr& A0, const TensorPtr& A1, const TensorPtr& X, const TensorPtr& Xinv,
            const BSTensor& XLY, double lead, TensorPtr& A0n, TensorPtr& A1n, std::vector<double>& lambda0n,
            OrthoInfo& info) {
  TensorPtr U, Vd;
  double disc = 0.0;
  svd_truncate(ctx, XLY, 0, 0.0, U, Vd, lambda0n, disc);
  double sq = 0.0;
  for (double v : lambda0n) sq += v * v;
  for (double& v : lambda0n) v *= std::sqrt(1.0 / sq);
  info.entropy = entropy(lambda0n);
  launch_scale(cplx{1.0 / std::sqrt(lead), 0.0}, X->data.get(), X->total, ctx.stream);
  A0n = reorder(ctx, *contract(ctx, *contract(ctx, *U, true, *X, false, {{0, 0}}), false, *A0, false, {{1, 1}}),
                false, {1, 0, 2}, 2);
  A1n = contract(ctx, *A1, false, *contract(ctx, *Xinv, false, *U, false, {{1, 0}}), false, {{2, 0}});
  HIP_CHECK(hipStreamSynchronize(ctx.stream));
}

}  // namespace

void orthogonalise(Ctx& ctx, const TensorPtr& A0, const TensorPtr& A1, TensorPtr& A0n, TensorPtr& A1n,
                   std::vector<double>& lambda0n, OrthoInfo& info) {
  const auto t0 = std::chrono::steady_clock::now();
  const std::vector<TensorPtr> cell{A0, A1};
  const double tol = 1e-14;
  TensorPtr ML, MR, X, Xinv, Y;
  info.left = transfer_leading(ctx, cell, 0, tol, &ML);
  check_leading(info.left, "left");
  info.kept_left = sqrt_dr(ctx, *ML, &X, &Xinv, nullptr);
  ML.reset();
  info.right = transfer_leading(ctx, cell, 1, tol, &MR);
  check_leading(info.right, "right");
  info.kept_right = sqrt_dr(ctx, *MR, nullptr, nullptr, &Y);
  MR.reset();
  TensorPtr XLY = contract(ctx, *X, false, *Y, false, {{1, 0}});
  finish(ctx, A0, A1, X, Xinv, *XLY, info.left.value.re, A0n, A1n, lambda0n, info);
  info.seconds = seconds_since(t0);
}

void orthogonalise_inversion_symmetric(Ctx& ctx, const TensorPtr& A0, const TensorPtr& A1,
                                       const std::vector<double>& lambda0, TensorPtr& A0n, TensorPtr& A1n,
                                       std::vector<double>& lambda0n, OrthoInfo& info) {
  const auto t0 = std::chrono::steady_clock::now();
  const std::vector<TensorPtr> cell{A0, A1};
  TensorPtr ML, X, Xinv, Y;
  info.left = transfer_leading(ctx, cell, 0, 1e-14, &ML);
  check_leading(info.left, "left");
  info.right = info.left;
  // the left fixed point serves on both boundaries (the reference runs
  // SqrtDR twice on the same matrix; one decomposition gives X, X^-1, Y)
  info.kept_left = info.kept_right = sqrt_dr(ctx, *ML, &X, &Xinv, &Y);
  ML.reset();
  // X . diag(lambda0) . Y, lambda0 on the outward index of the last matrix
  TensorPtr Xl = reorder(ctx, *X, false, {0, 1}, 1);
  scale_along_index(ctx, *Xl, 1, lambda0);
  TensorPtr XLY = contract(ctx, *Xl, false, *Y, false, {{1, 0}});
  finish(ctx, A0, A1, X, Xinv, *XLY, info.left.value.re, A0n, A1n, lambda0n, info);
  info.seconds = seconds_since(t0);
}

cplx one_vertex_measurement(Ctx& ctx, const BSTensor& Op, const TensorPtr& A0, const TensorPtr& A1,
                            const std::vector<double>& lambda0) {
  CAM_REQUIRE(Op.nidx() == 4, "OneVertexMeasurement: the operator must be an MPO (four indices)");
  std::vector<double> l2(lambda0.size());
  for (size_t k = 0; k < l2.size(); ++k) l2[k] = lambda0[k] * lambda0[k];
  const TensorPtr L2 = diagonal_tensor(ctx, A0->idx[1], l2, A0->rules);
  TensorPtr t = contract(ctx, *A0, true, *A0, false, {{0, 0}, {1, 1}});
  t = contract(ctx, *t, false, *A1, false, {{1, 1}});
  t = contract(ctx, *t, false, Op, false, {{1, 2}});
  t = contract(ctx, *A1, true, *t, false, {{0, 2}, {1, 0}});
  return scalar_of(ctx, *contract(ctx, *t, false, *L2, false, {{0, 1}, {1, 0}}));
}

cplx i_two_vertex_energy(Ctx& ctx, const BSTensor& W, const TensorPtr& A0, const TensorPtr& A1,
                         const std::vector<double>& lambda0, cplx terms[3]) {
  CAM_REQUIRE(W.nidx() == 4, "iTwoVertexEnergy: the Hamiltonian must be an MPO (four indices)");
  const int64_t w = W.idx[1].size;
  CAM_REQUIRE(w >= 2 && W.idx[3].size == w, "iTwoVertexEnergy: malformed MPO");
  auto c = [&](const TensorPtr& a, bool ca, const TensorPtr& b, bool cb, std::vector<std::pair<int, int>> p) {
    return contract(ctx, *a, ca, *b, cb, p);
  };
  std::vector<double> l2(lambda0.size());
  for (size_t k = 0; k < l2.size(); ++k) l2[k] = lambda0[k] * lambda0[k];
  // contract(LAMBDA0, LAMBDA0, contract10): diag(lambda0^2) on A0's inward index
  const TensorPtr L2 = diagonal_tensor(ctx, A0->idx[1], l2, A0->rules);
  auto close = [&](const TensorPtr& t, std::vector<std::pair<int, int>> p) {
    return to_z(scalar_of(ctx, *c(t, false, L2, false, p)));
  };
  zcplx local = 0.0, bond1 = 0.0, bond2 = 0.0;
  const TensorPtr H1 = extract_mpo_block(ctx, W, w - 1, w - 1, 0, 0);
  if (tensor_to_csc(*H1, 0, nullptr, nullptr, nullptr, ctx.stream) > 0) {
    // H1 on the first vertex
    TensorPtr t = c(H1, false, A0, false, {{2, 0}});
    t = c(A0, true, t, false, {{0, 0}, {1, 3}});
    t = c(t, false, A1, false, {{3, 1}});
    t = c(A1, true, t, false, {{0, 3}, {1, 0}});
    local += close(t, {{0, 1}, {3, 0}});
    // OneVertexMeasurement: H1 on the second vertex
    local += to_z(one_vertex_measurement(ctx, *H1, A0, A1, lambda0));
  }
  if (w >= 3) {
    const TensorPtr ColX = extract_mpo_block(ctx, W, 1, w - 2, 0, 0);
    const TensorPtr RowX = extract_mpo_block(ctx, W, w - 1, w - 1, 1, W.idx[3].size - 2);
    // TwoVertexMeasurement(RowX, ColX, A0, A1): the bond inside the cell
    TensorPtr t = c(A0, true, A0, false, {{1, 1}});
    t = c(t, false, RowX, false, {{0, 0}, {2, 2}});
    t = c(t, false, A1, true, {{0, 1}});
    t = c(t, false, ColX, false, {{2, 1}, {3, 0}});
    t = c(t, false, A1, false, {{0, 1}, {3, 0}});
    bond1 = close(t, {{1, 1}, {3, 0}});
    // the bond from the second vertex to the first vertex of the next cell
    t = c(A0, true, A0, false, {{0, 0}, {1, 1}});
    t = c(t, false, A1, false, {{1, 1}});
    t = c(t, false, RowX, false, {{1, 2}});
    t = c(A1, true, t, false, {{0, 2}, {1, 0}});
    t = c(A0, false, t, false, {{1, 1}});
    t = c(t, false, ColX, false, {{0, 2}, {4, 1}});
    t = c(A0, true, t, false, {{0, 3}, {1, 1}});
    t = c(t, false, A1, false, {{1, 1}});
    t = c(A1, true, t, false, {{0, 3}, {1, 0}});
    bond2 = close(t, {{0, 1}, {3, 0}});
  }
  if (terms) {
    terms[0] = to_c(local);
    terms[1] = to_c(bond1);
    terms[2] = to_c(bond2);
  }
  return to_c(0.5 * (local + bond1 + bond2));
}

}  // namespace cam
