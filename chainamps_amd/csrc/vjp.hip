// Adjoint (vector-Jacobian product) of the block-sparse H_eff contraction.
//
// No counterpart in the reference.  For E = <psi|H_eff|psi>/<psi|psi> with
// fixed environments L, R and the MPO W entering both vertices:
//   dE/dconj(psi) = (H psi - E psi) / <psi|psi>
//   dE/dW         = [ sum over the W of vertex 1 + the W of vertex 2 ] / <psi|psi>
// computed by back-propagating conj(psi) through the four matvec stages:
//   Tb3[bR](a0',a2,s1',s2') = sum_{a2'} conj(psi)(a0',a2',s1',s2') R_bR(a2',a2)
//   Tb2[b1](a0',a2,s1',s2)  = sum_{bR,s2'} Tb3[bR] W(s2',b1,s2,bR)
//   dW(s1',bL,s1,b1) += sum_{a0',a2,s2}  Tb2[b1](..,s1',s2) T1[bL](..,s1,s2)
//   dW(s2',b1,s2,bR) += sum_{a0',a2,s1'} Tb3[bR](..,s2')   T2[b1](..,s2)
// T1, T2 are the forward intermediates of one FP64 matvec (environments
// and intermediates are recomputed per call, never stored per Lanczos
// step).  The W reductions have small M, N (sigma sectors) and long K, so
// K is split into chunks written to partial slots and summed by a second
// kernel.
#include <algorithm>
#include <map>

#include "solvers.hpp"

namespace cam {

namespace {

struct RedJob {
  const double2* slots;   // nslot consecutive M x N column-major partials
  int64_t nslot, M, N;
  double2* out;           // out[m + n*ld]
  int64_t ld;
  double scale;
};

__global__ void reduce_slots_kernel(const RedJob* jobs, int njobs) {
  for (int j = blockIdx.y; j < njobs; j += gridDim.y) {
    const RedJob J = jobs[j];
    const int64_t mn = J.M * J.N;
    for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < mn;
         e += int64_t(gridDim.x) * blockDim.x) {
      double re = 0.0, im = 0.0;
      for (int64_t s = 0; s < J.nslot; ++s) {
        const double2 v = J.slots[s * mn + e];
        re += v.x;
        im += v.y;
      }
      const int64_t m = e % J.M, n = e / J.M;
      J.out[m + n * J.ld] = make_double2(re * J.scale, im * J.scale);
    }
  }
}

constexpr int64_t kSplitK = 8192;

}  // namespace

void heff_vjp(Heff& H, const cplx* psi_ref, cplx* dpsi_ref, cplx* dW_out, int64_t nnzW) {
  hipStream_t s = H.ctx->stream;
  const int64_t n = H.n;
  H.build_perm();
  DevBuf<cplx> r(n), x(n), y(n);
  r.upload(psi_ref, n, s);
  launch_scatter_map(r.get(), x.get(), H.perm.get(), n, s);
  H.apply(x.get(), y.get(), 0);
  // the matvec's stage C wrote T3 over T1: recompute T1 and T2 of this psi
  // (the environments are recomputed, not stored per Lanczos step)
  H.forward_stages(x.get(), 2);
  DevBuf<cplx> partial(static_cast<size_t>(multi_dot_blocks(n)) * 64), coef(64);
  auto dot = [&](const cplx* a, const cplx* b) {
    launch_multi_dot(a, n, 1, b, n, partial.get(), coef.get(), s);
    cplx c;
    HIP_CHECK(hipMemcpyAsync(&c, coef.get(), sizeof(cplx), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return c;
  };
  const double n2 = dot(x.get(), x.get()).re;
  CAM_REQUIRE(n2 > 0.0, "vjp: zero vector");
  const double E = dot(x.get(), y.get()).re / n2;
  // dE/dconj(psi)
  launch_axpy(cplx{-E, 0.0}, x.get(), y.get(), n, s);
  launch_scale(cplx{1.0 / n2, 0.0}, y.get(), n, s);
  launch_gather_map(y.get(), r.get(), H.perm.get(), n, s);
  HIP_CHECK(hipMemcpyAsync(dpsi_ref, r.get(), n * sizeof(cplx), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));

  // ---- backward through stages D and C ------------------------------------
  DevBuf<cplx> Tb3(std::max<int64_t>(H.t3n, 1)), Tb2(std::max<int64_t>(H.t2n, 1));
  const cplx* Rb = H.R->data.get();
  const cplx* Wb = H.W->data.get();
  GemmPlan p3, p2, pw;
  for (auto& kv : H.t3) {
    const auto& k = kv.first;             // (bR, s0', s3, s1', s2')
    const Heff::TB& tb = kv.second;
    auto it = H.Rv.find({k[0], k[2]});
    if (it == H.Rv.end()) {
      launch_zero(Tb3.get() + tb.off, tb.m0 * tb.m3 * tb.n1 * tb.n2 * sizeof(cplx), s);
      continue;
    }
    const EnvMat& e = it->second;
    const auto pit = H.pindex.find({k[1], e.sout, k[3], k[4]});
    CAM_REQUIRE(pit != H.pindex.end(), "vjp: missing psi block");
    const Heff::PB& P = H.pblocks[pit->second];
    const int pi = p3.add_problem(Tb3.get() + tb.off, tb.m0, static_cast<int>(tb.m0), static_cast<int>(tb.m3),
                                  static_cast<int>(tb.n1 * tb.n2), tb.m0 * tb.m3, 0);
    p3.add_segment(pi, x.get() + P.off, P.m0, P.m0 * P.m3, Rb + e.off, e.ld, 0, static_cast<int>(P.m3),
                   kSegConjA);
  }
  for (auto& kv : H.t2) {
    const auto& k = kv.first;             // (b1, s0', s3, s1', s2)
    const Heff::TB& tb = kv.second;
    const int64_t rows = tb.m0 * tb.m3 * tb.n1;
    int pi = -1;
    for (auto& wv : H.Wv) {
      const auto& wk = wv.first;          // (bl, br, s', s)
      if (wk[0] != k[0] || wk[3] != k[4]) continue;
      auto t3 = H.t3.find({static_cast<int>(wk[1]), k[1], k[2], k[3], static_cast<int>(wk[2])});
      if (t3 == H.t3.end()) continue;
      if (pi < 0)
        pi = p2.add_problem(Tb2.get() + tb.off, rows, static_cast<int>(rows), static_cast<int>(tb.n2), 1, 0, 0);
      p2.add_segment(pi, Tb3.get() + t3->second.off, rows, 0, Wb + wv.second.off, wv.second.ld, 0,
                     static_cast<int>(t3->second.n2), 0);
    }
    if (pi < 0) launch_zero(Tb2.get() + tb.off, rows * tb.n2 * sizeof(cplx), s);
  }
  for (GemmPlan* p : {&p3, &p2}) {
    p->choose_tiles(1);
    p->build_tiles();
  }
  launch_grouped_gemm(p3, 0, s);
  launch_grouped_gemm(p2, 0, s);

  // ---- W gradient: products grouped by W view, K split into slots -----------
  struct Prod { const cplx* A; int64_t csA; const cplx* B; int64_t ldB; int64_t K; };
  std::map<std::array<int64_t, 4>, std::vector<Prod>> prods;
  for (auto& kv : H.t2) {
    const auto& k2 = kv.first;            // (b1, s0', s3, s1', s2)
    const Heff::TB& t2 = kv.second;
    const int64_t rows = t2.m0 * t2.m3;
    for (auto& wv : H.Wv) {
      const auto& wk = wv.first;          // (bL, b1, s1', s1)
      if (wk[1] != k2[0] || wk[2] != k2[3]) continue;
      auto t1 = H.t1.find({static_cast<int>(wk[0]), k2[1], k2[2], static_cast<int>(wk[3]), k2[4]});
      if (t1 == H.t1.end()) continue;
      for (int64_t q = 0; q < t2.n2; ++q)
        prods[wk].push_back(Prod{Tb2.get() + t2.off + q * rows * t2.n1, rows,
                                 H.T1.get() + t1->second.off + q * rows * t1->second.n1, rows, rows});
    }
  }
  for (auto& kv : H.t3) {
    const auto& k3 = kv.first;            // (bR, s0', s3, s1', s2')
    const Heff::TB& t3 = kv.second;
    const int64_t rows = t3.m0 * t3.m3 * t3.n1;
    for (auto& wv : H.Wv) {
      const auto& wk = wv.first;          // (b1, bR, s2', s2)
      if (wk[1] != k3[0] || wk[2] != k3[4]) continue;
      auto t2 = H.t2.find({static_cast<int>(wk[0]), k3[1], k3[2], k3[3], static_cast<int>(wk[3])});
      if (t2 == H.t2.end()) continue;
      prods[wk].push_back(Prod{Tb3.get() + t3.off, rows, H.T2.get() + t2->second.off, rows, rows});
    }
  }
  // slots
  int64_t nslots_tot = 0;
  struct KeySlots { int64_t first, count, M, N; WMat w; };
  std::vector<KeySlots> ks;
  for (auto& kv : prods) {
    const WMat& w = H.Wv.at(kv.first);
    KeySlots k{nslots_tot, 0, w.nout, w.nin, w};
    for (auto& p : kv.second) k.count += (p.K + kSplitK - 1) / kSplitK;
    nslots_tot += k.count * w.nout * w.nin;
    ks.push_back(k);
  }
  DevBuf<cplx> slots(std::max<int64_t>(nslots_tot, 1));
  {
    size_t ki = 0;
    for (auto& kv : prods) {
      const KeySlots& k = ks[ki++];
      int64_t slot = 0;
      for (auto& p : kv.second)
        for (int64_t k0 = 0; k0 < p.K; k0 += kSplitK) {
          const int64_t kk = std::min(kSplitK, p.K - k0);
          cplx* C = slots.get() + k.first + slot * k.M * k.N;
          const int pi = pw.add_problem(C, k.M, static_cast<int>(k.M), static_cast<int>(k.N), 1, 0, 0);
          pw.add_segment(pi, p.A + k0, p.csA, 0, p.B + k0, p.ldB, 0, static_cast<int>(kk), kSegAKContig);
          ++slot;
        }
    }
  }
  pw.choose_tiles(1);
  pw.build_tiles();
  launch_grouped_gemm(pw, 0, s);
  DevBuf<cplx> gW(std::max<int64_t>(H.W->total, 1));
  launch_zero(gW.get(), std::max<int64_t>(H.W->total, 1) * sizeof(cplx), s);
  std::vector<RedJob> jobs;
  for (auto& k : ks)
    jobs.push_back(RedJob{reinterpret_cast<const double2*>(slots.get() + k.first), k.count, k.M, k.N,
                          reinterpret_cast<double2*>(gW.get() + k.w.off), k.w.ld, 1.0 / n2});
  if (!jobs.empty()) {
    DevBuf<RedJob> dj(jobs.size());
    dj.upload(jobs.data(), jobs.size(), s);
    hipLaunchKernelGGL(reduce_slots_kernel, dim3(16, std::min<size_t>(jobs.size(), 1024)), dim3(256), 0, s,
                       dj.get(), static_cast<int>(jobs.size()));
    HIP_CHECK(hipGetLastError());
  }
  HIP_CHECK(hipStreamSynchronize(s));

  // ---- values at W's stored (non-zero) entries, CSC order --------------------
  std::vector<cplx> hw = download(*H.W, s), hg(static_cast<size_t>(H.W->total));
  HIP_CHECK(hipMemcpy(hg.data(), gW.get(), H.W->total * sizeof(cplx), hipMemcpyDeviceToHost));
  const BSTensor& W = *H.W;
  const int nd = W.nidx();
  std::vector<int64_t> rst(nd, 0), cst(nd, 0);
  {
    int64_t m = 1;
    for (int k = 0; k < W.numrows; ++k) { rst[k] = m; m *= W.idx[k].size; }
    m = 1;
    for (int k = W.numrows; k < nd; ++k) { cst[k] = m; m *= W.idx[k].size; }
  }
  struct Ent { int64_t col, row; cplx g; };
  std::vector<Ent> ents;
  for (const Block& bl : W.blocks)
    for (int64_t e = 0; e < bl.size; ++e) {
      const cplx v = hw[bl.offset + e];
      if (v.re == 0.0 && v.im == 0.0) continue;
      int64_t rem = e, row = 0, col = 0;
      for (int k = 0; k < nd; ++k) {
        const int64_t o = rem % bl.dims[k];
        rem /= bl.dims[k];
        const int64_t pos = W.idx[k].sec_pos[bl.sec[k]][o];
        if (k < W.numrows) row += pos * rst[k]; else col += pos * cst[k];
      }
      ents.push_back({col, row, hg[bl.offset + e]});
    }
  std::sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) {
    return a.col != b.col ? a.col < b.col : a.row < b.row;
  });
  CAM_REQUIRE(static_cast<int64_t>(ents.size()) == nnzW, "vjp: nnzW does not match the MPO's stored entries");
  for (size_t i = 0; i < ents.size(); ++i) dW_out[i] = ents[i].g;
}

}  // namespace cam
