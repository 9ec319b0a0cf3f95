// Blockwise SVD with the reference truncation rule, FP64 on the device.
//
// MPX_matrix::SVD (MPX_matrix.cpp:157-192) -> SparseMatrix::SVD
// (sparse_interface.cpp:1248-1406):
//   * columns are grouped by OutgoingColState (GetAllBlockColumns,
//     MPX_matrix.cpp:331-355); in block storage a group is the set of
//     blocks whose column sectors carry that charge;
//   * a group whose squared weight is <= eps^2 is skipped;
//   * the dense block keeps only non-empty rows and columns
//     (TranslationBlock, sparse_interface.hpp:353-371) and is decomposed
//     by economy SVD (the reference calls zgesdd with JOB='S',
//     dense_matrix_functions.cpp:276-337; here rocSOLVER zgesvd);
//   * all singular values are sorted descending (ties: block order, then
//     index inside the block -- std::sort leaves ties unspecified), the
//     kept length is min(D, #values) (D=0: all), cut further while the
//     discarded squared tail stays below max_trunc, then extended through
//     relative gaps < 1e-3;
//   * U entries with |x| <= eps/rows and V^dagger entries <= eps/cols are
//     dropped; the new bond carries the ingoing row state of each U column
//     (MPX_matrix.cpp:176-180), i.e. the group charge.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <map>
#include <cstdlib>
#include <numeric>
#include <exception>
#include <thread>

#include "solvers.hpp"

namespace cam {

namespace {

// Blocks with min(rows, cols) >= this and room for at most half their values
// use the leading-values decomposition (CAM_SVD_EXACT=1 disables it).
constexpr int64_t kTruncatedMinDim = 1024;
// a call uses the second lane when at least two of its decompositions have
// this many elements (calls on tiny blocks stay on one lane and one thread)
constexpr int64_t kLaneMinElems = int64_t(1) << 14;
// Residual bound |A v - s u| of accepted triplets, relative to the block's
// Frobenius norm: each accepted value is within this of a singular value
// of the block.  The FP64 floor of the residual itself grows like
// eps sqrt(m + n) |A| (a few 1e-15 absolute for unit-norm blocks of 10^4
// rows), so larger blocks get max(5e-14, 2 eps sqrt(m + n)).
constexpr double kRitzTol = 5e-14;
// CAM_SVD_SERIAL=1 runs every decomposition of a call one after the other
// on the context's stream (no second lane)
bool serial_lanes() {
  static const bool serial = [] {
    const char* e = std::getenv("CAM_SVD_SERIAL");
    return e && *e && *e != '0';
  }();
  return serial;
}

double ritz_tol(int64_t m, int64_t n, double weight) {
  const double floor = 2.0 * std::numeric_limits<double>::epsilon() * std::sqrt(double(m + n));
  return std::max(kRitzTol, floor) * std::sqrt(weight);
}

struct FlagJob {
  const double2* in;
  int64_t M, N;          // block viewed as (rows x cols) column-major
  int* rflag;
  int* cflag;
};

__global__ void nonzero_flags_kernel(const FlagJob* jobs, int njobs) {
  for (int j = blockIdx.y; j < njobs; j += gridDim.y) {
    const FlagJob J = jobs[j];
    const int64_t tot = J.M * J.N;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < tot;
         e += int64_t(gridDim.x) * blockDim.x) {
      const double2 v = J.in[e];
      if (v.x != 0.0 || v.y != 0.0) {
        J.rflag[e % J.M] = 1;
        J.cflag[e / J.M] = 1;
      }
    }
  }
}

// out[i + j*ldo] = (r[i] >= 0 && c[j] >= 0) ? in[r[i]*rs + c[j]*cs] : 0,
// with |x| <= thr dropped to 0
struct GatherJob {
  const double2* in;
  double2* out;
  const int64_t* r;
  const int64_t* c;
  int64_t M, N, ldo, rs, cs;
  double thr;
};

__global__ void gather2d_kernel(const GatherJob* jobs, int njobs) {
  for (int j = blockIdx.y; j < njobs; j += gridDim.y) {
    const GatherJob J = jobs[j];
    const int64_t tot = J.M * J.N;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < tot;
         e += int64_t(gridDim.x) * blockDim.x) {
      const int64_t i = e % J.M, jj = e / J.M;
      const int64_t ri = J.r[i], cj = J.c[jj];
      double2 v = make_double2(0.0, 0.0);
      if (ri >= 0 && cj >= 0) {
        v = J.in[ri * J.rs + cj * J.cs];
        if (v.x * v.x + v.y * v.y <= J.thr * J.thr) v = make_double2(0.0, 0.0);
      }
      J.out[i + jj * J.ldo] = v;
    }
  }
}

template <typename J, typename K>
void run_jobs(K kernel, const std::vector<J>& jobs, hipStream_t s) {
  if (jobs.empty()) return;
  DevBuf<J> d(jobs.size());
  d.upload(jobs.data(), jobs.size(), s);
  const int gy = static_cast<int>(std::min<size_t>(jobs.size(), 1024));
  hipLaunchKernelGGL(kernel, dim3(64, gy), dim3(256), 0, s, d.get(), static_cast<int>(jobs.size()));
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(s));
}

struct Tuple {
  std::vector<int32_t> sec;
  std::vector<int64_t> dims;
  int64_t size = 1;
  bool operator<(const Tuple& o) const { return sec < o.sec; }
};

struct Group {
  Charge q;
  std::vector<Tuple> rows, cols;                 // distinct sector tuples
  std::map<std::vector<int32_t>, int> rpos, cpos;
  std::vector<int64_t> roff, coff;               // uncompacted offsets of tuples
  int64_t rtot = 0, ctot = 0;
  std::vector<int> blocks;
  // compaction
  std::vector<int64_t> rmap, cmap;               // uncompacted -> compact (-1 empty)
  int64_t mg = 0, ng = 0;
  int64_t first_col = 0;                         // GetAllBlockColumns order key
  double weight = 0.0;
  int64_t k = 0;                                 // number of singular values
  DevBuf<cplx> A, U, Vh;
  std::vector<double> S;
  bool truncated = false;                        // only the leading k values computed
  double tail = 0.0;                             // weight - sum S^2 when truncated
  double last_res = 0.0;                         // residual of the last computed triplet
  int64_t kk = 0;                                // leading triplets asked for (truncated path)
  int64_t nver = 0;                              // leading triplets verified at acceptance
  std::vector<double> res;                       // residuals of the first ncheck triplets
};

}  // namespace

namespace {

void svd_core(Ctx& ctx, const BSTensor& th, int64_t D, double max_trunc, TensorPtr& Uout, TensorPtr& Vdout,
              std::vector<double>& values, double& discarded, bool values_only) {
  hipStream_t s = ctx.stream;
  const double eps = std::numeric_limits<double>::epsilon();
  const int r = th.numrows, N = th.nidx();
  CAM_REQUIRE(r >= 1 && r < N, "SVD needs at least one row and one column index");
  // ---- groups of blocks by column state -----------------------------------
  std::map<Charge, Group> groups;
  for (size_t b = 0; b < th.blocks.size(); ++b) {
    const Block& bl = th.blocks[b];
    Charge qc = th.rules.zero();
    for (int k = r; k < N; ++k) {
      const Charge& q = th.idx[k].sec_q[bl.sec[k]];
      qc = th.idx[k].inward ? th.rules.sub(qc, q) : th.rules.add(qc, q);
    }
    Group& g = groups[qc];
    g.q = qc;
    g.blocks.push_back(static_cast<int>(b));
    Tuple rt, ct;
    rt.sec.assign(bl.sec.begin(), bl.sec.begin() + r);
    rt.dims.assign(bl.dims.begin(), bl.dims.begin() + r);
    ct.sec.assign(bl.sec.begin() + r, bl.sec.end());
    ct.dims.assign(bl.dims.begin() + r, bl.dims.end());
    for (auto d : rt.dims) rt.size *= d;
    for (auto d : ct.dims) ct.size *= d;
    if (!g.rpos.count(rt.sec)) { g.rpos[rt.sec] = -1; g.rows.push_back(rt); }
    if (!g.cpos.count(ct.sec)) { g.cpos[ct.sec] = -1; g.cols.push_back(ct); }
  }
  // column strides of the flat column index (for the group order)
  std::vector<int64_t> cstride(N, 0);
  {
    int64_t m = 1;
    for (int k = r; k < N; ++k) { cstride[k] = m; m *= th.idx[k].size; }
  }
  std::vector<Group*> glist;
  for (auto& kv : groups) {
    Group& g = kv.second;
    std::sort(g.rows.begin(), g.rows.end());
    std::sort(g.cols.begin(), g.cols.end());
    for (size_t i = 0; i < g.rows.size(); ++i) {
      g.rpos[g.rows[i].sec] = static_cast<int>(i);
      g.roff.push_back(g.rtot);
      g.rtot += g.rows[i].size;
    }
    g.first_col = std::numeric_limits<int64_t>::max();
    for (size_t i = 0; i < g.cols.size(); ++i) {
      g.cpos[g.cols[i].sec] = static_cast<int>(i);
      g.coff.push_back(g.ctot);
      g.ctot += g.cols[i].size;
      int64_t fc = 0;
      for (int k = r; k < N; ++k) fc += th.idx[k].sec_pos[g.cols[i].sec[k - r]][0] * cstride[k];
      g.first_col = std::min(g.first_col, fc);
    }
    glist.push_back(&g);
  }
  std::sort(glist.begin(), glist.end(), [](const Group* a, const Group* b) { return a->first_col < b->first_col; });

  // ---- non-empty rows / columns ---------------------------------------------
  int64_t ftot = 0;
  std::vector<int64_t> fbase(glist.size());
  for (size_t gi = 0; gi < glist.size(); ++gi) {
    fbase[gi] = ftot;
    ftot += glist[gi]->rtot + glist[gi]->ctot;
  }
  DevBuf<int> dflags(std::max<int64_t>(ftot, 1));
  launch_zero(dflags.get(), std::max<int64_t>(ftot, 1) * sizeof(int), s);
  {
    std::vector<FlagJob> jobs;
    for (size_t gi = 0; gi < glist.size(); ++gi) {
      Group& g = *glist[gi];
      for (int b : g.blocks) {
        const Block& bl = th.blocks[b];
        std::vector<int32_t> rs(bl.sec.begin(), bl.sec.begin() + r), cs(bl.sec.begin() + r, bl.sec.end());
        const int ri = g.rpos[rs], ci = g.cpos[cs];
        FlagJob j{reinterpret_cast<const double2*>(th.data.get() + bl.offset), g.rows[ri].size,
                  g.cols[ci].size, dflags.get() + fbase[gi] + g.roff[ri],
                  dflags.get() + fbase[gi] + g.rtot + g.coff[ci]};
        if (j.M * j.N) jobs.push_back(j);
      }
    }
    run_jobs(nonzero_flags_kernel, jobs, s);
  }
  std::vector<int> flags(std::max<int64_t>(ftot, 1));
  HIP_CHECK(hipMemcpy(flags.data(), dflags.get(), std::max<int64_t>(ftot, 1) * sizeof(int), hipMemcpyDeviceToHost));
  for (size_t gi = 0; gi < glist.size(); ++gi) {
    Group& g = *glist[gi];
    g.rmap.assign(g.rtot, -1);
    g.cmap.assign(g.ctot, -1);
    for (int64_t i = 0; i < g.rtot; ++i)
      if (flags[fbase[gi] + i]) g.rmap[i] = g.mg++;
    for (int64_t j = 0; j < g.ctot; ++j)
      if (flags[fbase[gi] + g.rtot + j]) g.cmap[j] = g.ng++;
  }

  // ---- assemble dense group matrices and decompose -----------------------
  DevBuf<cplx> partial(static_cast<size_t>(multi_dot_blocks(int64_t(1) << 40))), dcoef(64);   // k = 1 below
  const char* ev = std::getenv("CAM_SVD_EXACT");
  const bool exact_only = ev && *ev && *ev != '0';
  // CAM_SVD_TRACE=1 prints the path and time of every decomposed block
  const char* tv = std::getenv("CAM_SVD_TRACE");
  const bool trace = tv && *tv && *tv != '0';
  // per-lane result buffers of the full decompositions (grown before the
  // lanes start, so that no lane frees device memory while the other runs)
  struct LaneBufs {
    DevBuf<double> dS, dE;
    DevBuf<int> dinfo;
  };
  LaneBufs lane_bufs[2];
  for (LaneBufs& lb : lane_bufs) lb.dinfo.alloc(1);
  // the full (economy) decomposition of one group, as the reference's
  // per-block zgesdd, on the stream and rocBLAS handle of `lane`
  auto full_svd = [&](Ctx& lane, LaneBufs& lb, Group& g) {
    const hipStream_t ls = lane.stream;
    g.k = std::min(g.mg, g.ng);
    g.truncated = false;
    g.tail = 0.0;
    // values only: no singular vectors (U, V^dagger are 1-element placeholders)
    g.U.alloc(values_only ? 1 : g.mg * g.k);
    g.Vh.alloc(values_only ? 1 : g.k * g.ng);
    lb.dS.ensure(g.k);
    lb.dE.ensure(std::max<int64_t>(g.k, 1));
    const rocblas_svect sv = values_only ? rocblas_svect_none : rocblas_svect_singular;
    const rocblas_status st = rocsolver_zgesvd(
        static_cast<rocblas_handle>(lane.blas_handle()), sv, sv, static_cast<rocblas_int>(g.mg),
        static_cast<rocblas_int>(g.ng), reinterpret_cast<rocblas_double_complex*>(g.A.get()),
        static_cast<rocblas_int>(g.mg), lb.dS.get(), reinterpret_cast<rocblas_double_complex*>(g.U.get()),
        static_cast<rocblas_int>(values_only ? 1 : g.mg), reinterpret_cast<rocblas_double_complex*>(g.Vh.get()),
        static_cast<rocblas_int>(values_only ? 1 : g.k), lb.dE.get(), rocblas_outofplace, lb.dinfo.get());
    CAM_REQUIRE(st == rocblas_status_success, "rocsolver_zgesvd failed");
    g.S.resize(g.k);
    int info = 0;
    HIP_CHECK(hipMemcpyAsync(g.S.data(), lb.dS.get(), g.k * sizeof(double), hipMemcpyDeviceToHost, ls));
    HIP_CHECK(hipMemcpyAsync(&info, lb.dinfo.get(), sizeof(int), hipMemcpyDeviceToHost, ls));
    HIP_CHECK(hipStreamSynchronize(ls));
    CAM_REQUIRE(info == 0, "blockwise SVD did not converge");
    g.A.reset();
  };
  auto timed_full_svd = [&](Ctx& lane, LaneBufs& lb, Group& g) {
    const auto t1 = std::chrono::steady_clock::now();
    full_svd(lane, lb, g);
    if (trace)
      std::fprintf(stderr, "svd_truncate: block %lld x %lld, full decomposition, %.3f s\n",
                   static_cast<long long>(g.mg), static_cast<long long>(g.ng),
                   std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count());
  };
  // one decomposition to run: the leading values of a large block, or a
  // full decomposition
  struct Pending {
    Group* g;
    int64_t kk, ncheck, naccept;
    double tol;
    bool ok;
    bool full;
  };
  std::vector<Pending> pending;
  for (size_t gi = 0; gi < glist.size(); ++gi) {
    Group& g = *glist[gi];
    if (g.mg == 0 || g.ng == 0) continue;
    CAM_REQUIRE(g.mg < (int64_t(1) << 31) && g.ng < (int64_t(1) << 31), "SVD block too large");
    g.A.alloc(g.mg * g.ng);
    launch_zero(g.A.get(), g.mg * g.ng * sizeof(cplx), s);
    // scatter the blocks: gather2d with the output as the compact matrix
    std::vector<GatherJob> jobs;
    std::vector<std::vector<int64_t>> keep;
    for (int b : g.blocks) {
      const Block& bl = th.blocks[b];
      std::vector<int32_t> rs(bl.sec.begin(), bl.sec.begin() + r), cs(bl.sec.begin() + r, bl.sec.end());
      const int ri = g.rpos[rs], ci = g.cpos[cs];
      // iterate over compact rows/cols that this block covers
      std::vector<int64_t> rr, cc, orow, ocol;
      for (int64_t i = 0; i < g.rows[ri].size; ++i)
        if (g.rmap[g.roff[ri] + i] >= 0) { rr.push_back(i); orow.push_back(g.rmap[g.roff[ri] + i]); }
      for (int64_t j = 0; j < g.cols[ci].size; ++j)
        if (g.cmap[g.coff[ci] + j] >= 0) { cc.push_back(j); ocol.push_back(g.cmap[g.coff[ci] + j]); }
      if (rr.empty() || cc.empty()) continue;
      // compact rows/cols of one block are contiguous ranges in the group
      // matrix (tuples are laid out back to back): out = A + orow0 + ocol0*mg
      keep.push_back(rr);
      keep.push_back(cc);
      GatherJob j{};
      j.in = reinterpret_cast<const double2*>(th.data.get() + bl.offset);
      j.out = reinterpret_cast<double2*>(g.A.get() + orow.front() + ocol.front() * g.mg);
      j.M = static_cast<int64_t>(rr.size());
      j.N = static_cast<int64_t>(cc.size());
      j.ldo = g.mg;
      j.rs = 1;
      j.cs = g.rows[ri].size;
      j.thr = 0.0;
      jobs.push_back(j);
    }
    // upload the row/col maps
    std::vector<int64_t> flat;
    std::vector<size_t> at;
    for (auto& v : keep) { at.push_back(flat.size()); flat.insert(flat.end(), v.begin(), v.end()); }
    DevBuf<int64_t> dmaps(std::max<size_t>(flat.size(), 1));
    dmaps.upload(flat.data(), flat.size(), s);
    for (size_t q = 0; q < jobs.size(); ++q) {
      jobs[q].r = dmaps.get() + at[2 * q];
      jobs[q].c = dmaps.get() + at[2 * q + 1];
    }
    run_jobs(gather2d_kernel, jobs, s);
    // group weight (sum of squared moduli)
    launch_multi_dot(g.A.get(), g.mg * g.ng, 1, g.A.get(), g.mg * g.ng, partial.get(), dcoef.get(), s);
    cplx wsum;
    HIP_CHECK(hipMemcpyAsync(&wsum, dcoef.get(), sizeof(cplx), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    g.weight = wsum.re;
    if (g.weight <= eps * eps) continue;                  // empty block: skipped
    // a large block of which at most D values can be kept: leading values
    // only (randomized subspace iteration, accepted on its Ritz residuals)
    const int64_t kfull = std::min(g.mg, g.ng);
    const int64_t kk = D > 0 ? std::min(kfull, D + std::max<int64_t>(D / 2, 32)) : kfull;
    if (!exact_only && D > 0 && kfull >= kTruncatedMinDim && 2 * kk <= kfull) {
      // residuals of the first min(D + 32, kk) triplets are computed; the
      // ones that must be within the tolerance for acceptance are the first
      // (kept count of the last block of this shape and charge) + margin,
      // or all of them without that hint.  The truncation below verifies
      // that everything it keeps was within the tolerance, with a verified
      // value below the cut after it.
      const int64_t ncheck = std::min(D + 32, kk);
      int64_t naccept = ncheck;
      const auto hk = std::make_pair(std::array<int64_t, 3>{g.mg, g.ng, kk}, g.q);
      const auto hit = ctx.svd_kept_hint.find(hk);
      if (hit != ctx.svd_kept_hint.end())
        naccept = std::min(ncheck, std::max<int64_t>(32, hit->second + std::max<int64_t>(32, hit->second / 4)));
      pending.push_back(Pending{&g, kk, ncheck, naccept, ritz_tol(g.mg, g.ng, g.weight), false, false});
      continue;
    }
    pending.push_back(Pending{&g, 0, 0, 0, 0.0, false, true});
  }

  // ---- decompositions --------------------------------------------------------
  // Two lanes: the context's stream and rocBLAS handle, and an auxiliary
  // pair driven from a second host thread, pull the decompositions from one
  // queue, largest first.  The latency-bound parts of one (Cholesky
  // factorisations, the Jordan-Wielandt core, rocSOLVER's bidiagonal sweeps,
  // the host synchronisations of the residual checks) then overlap the other
  // one's GEMMs.  Both streams have the default priority.  Every group matrix
  // was completed on the main stream, which is synchronised above; each
  // decomposition touches its own group only.
  auto run_truncated = [&](Ctx& lane, Pending& pj) {
    Group& g = *pj.g;
    std::vector<double> S;
    int rounds = 0;
    const auto t0 = std::chrono::steady_clock::now();
    pj.ok = truncated_svd(lane, g.A.get(), g.mg, g.ng, pj.kk, pj.ncheck, pj.tol, 24, S, g.U, g.Vh, &rounds,
                          &g.last_res, pj.naccept, &g.res);
    HIP_CHECK(hipStreamSynchronize(lane.stream));
    if (trace)
      std::fprintf(stderr, "svd_truncate: block %lld x %lld, leading %lld (first %lld verified): %s after %d "
                   "rounds, %.3f s\n",
                   static_cast<long long>(g.mg), static_cast<long long>(g.ng), static_cast<long long>(pj.kk),
                   static_cast<long long>(pj.naccept), pj.ok ? "accepted" : "not accepted", rounds,
                   std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    if (pj.ok) {
      g.S = std::move(S);
      g.k = pj.kk;
      g.kk = pj.kk;
      g.nver = pj.naccept;
      g.truncated = true;
      double kept = 0.0;
      for (double v : g.S) kept += v * v;
      g.tail = std::max(0.0, g.weight - kept);
    }
  };
  auto run_item = [&](Ctx& lane, LaneBufs& lb, Pending& pj) {
    if (pj.full) timed_full_svd(lane, lb, *pj.g);
    else run_truncated(lane, pj);
  };
  // cost order: the leading values of a block take about 8 k m n (four
  // rounds of two products), a full decomposition about min(m, n) m n
  auto cost = [](const Pending& pj) {
    const double mn = double(pj.g->mg) * double(pj.g->ng);
    return pj.full ? mn * double(std::min(pj.g->mg, pj.g->ng)) : 8.0 * mn * double(pj.kk);
  };
  std::stable_sort(pending.begin(), pending.end(),
                   [&](const Pending& a, const Pending& b) { return cost(a) > cost(b); });
  int64_t kmax = 1, nbig = 0;
  for (const Pending& pj : pending) {
    if (pj.full) kmax = std::max(kmax, std::min(pj.g->mg, pj.g->ng));
    if (pj.g->mg * pj.g->ng >= kLaneMinElems) ++nbig;
  }
  for (LaneBufs& lb : lane_bufs) {
    lb.dS.ensure(kmax);
    lb.dE.ensure(kmax);
  }
  if (nbig >= 2 && !serial_lanes()) {
    Ctx& aux = ctx.aux_lane();
    std::atomic<size_t> next{0};
    auto drain = [&](Ctx& lane, LaneBufs& lb) {
      for (size_t i = next++; i < pending.size(); i = next++) run_item(lane, lb, pending[i]);
    };
    std::exception_ptr err;
    std::thread helper([&]() {
      try {
        HIP_CHECK(hipSetDevice(aux.device));
        drain(aux, lane_bufs[1]);
      } catch (...) {
        next = pending.size();   // the other lane stops after its current item
        err = std::current_exception();
      }
    });
    std::exception_ptr mine;
    try {
      drain(ctx, lane_bufs[0]);
    } catch (...) {
      next = pending.size();
      mine = std::current_exception();
    }
    helper.join();
    if (mine) std::rethrow_exception(mine);
    if (err) std::rethrow_exception(err);
  } else {
    for (Pending& pj : pending) run_item(ctx, lane_bufs[0], pj);
  }
  for (Pending& pj : pending) {
    if (pj.full || pj.ok) continue;   // g.A is kept in case the full decomposition is needed
    pj.g->kk = pj.kk;
    timed_full_svd(ctx, lane_bufs[0], *pj.g);
  }

  // ---- global sort and truncation rule --------------------------------------
  // A truncated group contributes its leading values; the weight of its
  // uncomputed tail enters the discarded weight, which is exactly the sum
  // the full rule would accumulate over those values as long as every one
  // of them lies below the kept set.  That holds when each truncated
  // group's smallest computed value, raised by that triplet's residual (the
  // value is only known to lie within it of a singular value), is below the
  // last kept value by more than the 1e-3 degeneracy gap; otherwise the
  // group is decomposed in full and the rule is applied again.
  struct SV { double v; int g; int64_t i; };
  std::vector<SV> all;
  size_t length = 0;
  for (int pass = 0;; ++pass) {
    all.clear();
    double tail = 0.0;
    for (size_t gi = 0; gi < glist.size(); ++gi) {
      for (int64_t i = 0; i < glist[gi]->k; ++i) all.push_back(SV{glist[gi]->S[i], static_cast<int>(gi), i});
      if (glist[gi]->truncated) tail += glist[gi]->tail;
    }
    CAM_REQUIRE(!all.empty(), "Error: No singular values. Empty blocks?");
    std::stable_sort(all.begin(), all.end(), [](const SV& a, const SV& b) {
      if (a.v != b.v) return a.v > b.v;
      if (a.g != b.g) return a.g < b.g;
      return a.i < b.i;
    });
    const size_t nall = all.size();
    length = (D > 0 && static_cast<size_t>(D) < nall) ? static_cast<size_t>(D) : nall;
    double trunc_error = tail;
    size_t tl = nall;
    while (trunc_error < max_trunc && tl > 0) {
      trunc_error += all[tl - 1].v * all[tl - 1].v;
      --tl;
    }
    ++tl;
    if (tl < length) length = tl;
    while (length < nall && (all[length - 1].v - all[length].v) / all[length - 1].v < 1.0e-3) ++length;
    discarded = tail;
    for (size_t i = length; i < nall; ++i) discarded += all[i].v * all[i].v;
    bool redo = false;
    const double last = all[length - 1].v;
    const double below = last * (1.0 - 2.0e-3);
    std::vector<int64_t> kept_of(glist.size(), 0);
    for (size_t i = 0; i < length; ++i) kept_of[all[i].g]++;
    for (size_t gi = 0; gi < glist.size(); ++gi) {
      Group* g = glist[gi];
      if (!g->truncated) continue;
      // Every value a group keeps must be among its verified triplets, and
      // its last verified value, raised by its residual, must lie below the
      // cut: the verified Ritz values are the group's leading singular
      // values, so every later one -- checked but not yet converged, or not
      // computed at all -- is no larger and stays below the cut too.
      const int64_t nres = static_cast<int64_t>(g->res.size());
      const double tol = ritz_tol(g->mg, g->ng, g->weight);
      bool safe = kept_of[gi] < g->nver;
      if (safe) safe = g->S[g->nver - 1] + g->res[g->nver - 1] < below;
      if (safe) continue;
      redo = true;
      // verify every checked triplet (as without a kept-count hint), or as
      // many as the truncation keeps when that is more (the degeneracy
      // extension can keep values past D)
      const int64_t nver = std::min(g->kk, std::max(nres, kept_of[gi] + 16));
      if (g->nver < nver) {
        std::vector<double> S;
        int rounds = 0;
        if (truncated_svd(ctx, g->A.get(), g->mg, g->ng, g->kk, nver, tol, 24, S, g->U, g->Vh, &rounds,
                          &g->last_res, nver, &g->res)) {
          g->S = std::move(S);
          g->nver = nver;
          double kept = 0.0;
          for (double v : g->S) kept += v * v;
          g->tail = std::max(0.0, g->weight - kept);
          if (trace)
            std::fprintf(stderr, "svd_truncate: block %lld x %lld re-verified on %lld checked triplets "
                         "(keeps %lld)\n", static_cast<long long>(g->mg), static_cast<long long>(g->ng),
                         static_cast<long long>(nver), static_cast<long long>(kept_of[gi]));
          continue;
        }
      }
      full_svd(ctx, lane_bufs[0], *g);   // each pass turns at least one group exact: at most 2 #groups passes
    }
    if (!redo) break;
  }
  // the kept counts steer the next block of the same shape and charge
  {
    std::vector<int64_t> kept_of(glist.size(), 0);
    for (size_t i = 0; i < length; ++i) kept_of[all[i].g]++;
    for (size_t gi = 0; gi < glist.size(); ++gi)
      if (glist[gi]->kk > 0)
        ctx.svd_kept_hint[std::make_pair(std::array<int64_t, 3>{glist[gi]->mg, glist[gi]->ng, glist[gi]->kk},
                                         glist[gi]->q)] = kept_of[gi];
  }
  for (Group* g : glist) g->A.reset();
  values.clear();
  for (size_t i = 0; i < length; ++i) values.push_back(all[i].v);
  if (values_only) return;

  // ---- new bond index and the U / V^dagger tensors ------------------------
  QIndex nb;
  nb.physical = false;
  nb.size = static_cast<int64_t>(length);
  for (size_t i = 0; i < length; ++i) nb.charges.push_back(glist[all[i].g]->q);
  nb.build_sectors();
  std::vector<std::vector<int64_t>> kept(glist.size());     // local singular indices, sorted order
  for (size_t i = 0; i < length; ++i) kept[all[i].g].push_back(all[i].i);

  auto U = std::make_shared<BSTensor>();
  U->rules = th.rules;
  for (int k = 0; k < r; ++k) U->idx.push_back(th.idx[k]);
  QIndex nbo = nb;
  nbo.inward = false;
  U->idx.push_back(nbo);
  U->numrows = r;
  auto Vd = std::make_shared<BSTensor>();
  Vd->rules = th.rules;
  QIndex nbi = nb;
  nbi.inward = true;
  Vd->idx.push_back(nbi);
  for (int k = r; k < N; ++k) Vd->idx.push_back(th.idx[k]);
  Vd->numrows = 1;
  struct UB { int g; int tup; };
  std::vector<UB> ub, vb;
  for (size_t gi = 0; gi < glist.size(); ++gi) {
    Group& g = *glist[gi];
    if (kept[gi].empty()) continue;
    const int sq = nb.find_sector(g.q);
    for (size_t t = 0; t < g.rows.size(); ++t) {
      Block b;
      b.sec = g.rows[t].sec;
      b.sec.push_back(sq);
      b.dims = g.rows[t].dims;
      b.dims.push_back(static_cast<int64_t>(kept[gi].size()));
      U->blocks.push_back(b);
      ub.push_back(UB{static_cast<int>(gi), static_cast<int>(t)});
    }
    for (size_t t = 0; t < g.cols.size(); ++t) {
      Block b;
      b.sec.push_back(sq);
      b.sec.insert(b.sec.end(), g.cols[t].sec.begin(), g.cols[t].sec.end());
      b.dims.push_back(static_cast<int64_t>(kept[gi].size()));
      b.dims.insert(b.dims.end(), g.cols[t].dims.begin(), g.cols[t].dims.end());
      Vd->blocks.push_back(b);
      vb.push_back(UB{static_cast<int>(gi), static_cast<int>(t)});
    }
  }
  // sort blocks by code but keep the job association
  auto layout = [](BSTensor& T, std::vector<UB>& tag) {
    std::vector<size_t> ord(T.blocks.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::vector<uint64_t> codes(T.blocks.size());
    for (size_t i = 0; i < T.blocks.size(); ++i) codes[i] = T.code(T.blocks[i].sec);
    std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return codes[a] < codes[b]; });
    std::vector<Block> nbk;
    std::vector<UB> nt;
    for (auto o : ord) { nbk.push_back(T.blocks[o]); nt.push_back(tag[o]); }
    T.blocks.swap(nbk);
    tag.swap(nt);
    T.finalize_layout();
  };
  layout(*U, ub);
  layout(*Vd, vb);
  {
    std::vector<GatherJob> jobs;
    std::vector<int64_t> flat;
    std::vector<std::pair<size_t, size_t>> at;
    for (size_t bi = 0; bi < U->blocks.size(); ++bi) {
      Group& g = *glist[ub[bi].g];
      const int t = ub[bi].tup;
      size_t a = flat.size();
      for (int64_t i = 0; i < g.rows[t].size; ++i) flat.push_back(g.rmap[g.roff[t] + i]);
      size_t c = flat.size();
      for (auto li : kept[ub[bi].g]) flat.push_back(li);
      at.push_back({a, c});
      GatherJob j{};
      j.in = reinterpret_cast<const double2*>(g.U.get());
      j.out = reinterpret_cast<double2*>(U->data.get() + U->blocks[bi].offset);
      j.M = g.rows[t].size;
      j.N = static_cast<int64_t>(kept[ub[bi].g].size());
      j.ldo = j.M;
      j.rs = 1;
      j.cs = g.mg;
      j.thr = eps / static_cast<double>(g.mg);
      jobs.push_back(j);
    }
    size_t ubase = jobs.size();
    for (size_t bi = 0; bi < Vd->blocks.size(); ++bi) {
      Group& g = *glist[vb[bi].g];
      const int t = vb[bi].tup;
      size_t a = flat.size();
      for (auto li : kept[vb[bi].g]) flat.push_back(li);
      size_t c = flat.size();
      for (int64_t j = 0; j < g.cols[t].size; ++j) flat.push_back(g.cmap[g.coff[t] + j]);
      at.push_back({a, c});
      GatherJob j{};
      j.in = reinterpret_cast<const double2*>(g.Vh.get());
      j.out = reinterpret_cast<double2*>(Vd->data.get() + Vd->blocks[bi].offset);
      j.M = static_cast<int64_t>(kept[vb[bi].g].size());
      j.N = g.cols[t].size;
      j.ldo = j.M;
      j.rs = 1;
      j.cs = g.k;
      j.thr = eps / static_cast<double>(g.ng);
      jobs.push_back(j);
    }
    (void)ubase;
    DevBuf<int64_t> dmaps(std::max<size_t>(flat.size(), 1));
    dmaps.upload(flat.data(), flat.size(), s);
    for (size_t q = 0; q < jobs.size(); ++q) {
      jobs[q].r = dmaps.get() + at[q].first;
      jobs[q].c = dmaps.get() + at[q].second;
    }
    run_jobs(gather2d_kernel, jobs, s);
  }
  Uout = U;
  Vdout = Vd;
}

}  // namespace

void svd_truncate(Ctx& ctx, const BSTensor& th, int64_t D, double max_trunc, TensorPtr& Uout,
                  TensorPtr& Vdout, std::vector<double>& values, double& discarded) {
  svd_core(ctx, th, D, max_trunc, Uout, Vdout, values, discarded, false);
}

void singular_values(Ctx& ctx, const BSTensor& th, std::vector<double>& values) {
  TensorPtr U, Vd;
  double disc = 0.0;
  svd_core(ctx, th, 0, 0.0, U, Vd, values, disc, true);
}

}  // namespace cam
