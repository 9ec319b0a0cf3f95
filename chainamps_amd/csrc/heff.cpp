// H_eff planner and matvec (see heff.hpp for the stage definitions).
//
// Input conventions are those of the reference's TwoVertexComponents
// (DMRG_routines.cpp:779-833): the left block L has indices
// [dummy(out), a0'(in), dummy(in), bL(out), dummy(in), a0(out)], the MPO W
// [sigma'(in), bl(in), sigma(out), br(out)] and the right block R
// [a2'(out), dummy(in), bR(in), dummy(out), a2(in), dummy(out)].  In block
// storage the size-1 dummies vanish from the layout, so the L block of
// sectors (s0', sb, s0) is the array L[a0', b, a0] (a0' fastest) and the
// matrix L_b(a0', a0) for one bond value b is a strided view of it.
#include "heff.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace cam {

namespace {

// (b value, input sector) -> view; input sector = s0 (L) or s3 (R).
std::map<std::pair<int64_t, int>, EnvMat> env_views(const BSTensor& T, int iout, int ib, int iin,
                                                    const std::vector<cplx>& host, bool have_host) {
  std::map<std::pair<int64_t, int>, EnvMat> out;
  for (const Block& bl : T.blocks) {
    const int so = bl.sec[iout], sb = bl.sec[ib], si = bl.sec[iin];
    const int64_t mo = bl.dims[iout], nb = bl.dims[ib], mi = bl.dims[iin];
    const auto& bpos = T.idx[ib].sec_pos[sb];
    for (int64_t j = 0; j < nb; ++j) {
      // skip exactly-zero sub-blocks (e.g. the dummy left block selects one b)
      if (have_host) {
        bool nz = false;
        for (int64_t k = 0; k < mi && !nz; ++k)
          for (int64_t i = 0; i < mo; ++i) {
            const cplx v = host[bl.offset + i + mo * (j + nb * k)];
            if (v.re != 0.0 || v.im != 0.0) { nz = true; break; }
          }
        if (!nz) continue;
      }
      EnvMat e{so, bl.offset + mo * j, mo * nb, mo, mi};
      out[{bpos[j], si}] = e;
    }
  }
  return out;
}

}  // namespace

void Heff::build(Ctx& c, TensorPtr L_, TensorPtr W_, TensorPtr R_, double workspace_limit) {
  ctx = &c;
  L = L_;
  W = W_;
  R = R_;
  CAM_REQUIRE(L->nidx() == 6 && R->nidx() == 6 && W->nidx() == 4,
              "heff: expected 6-index blocks and a 4-index MPO");
  CAM_REQUIRE(L->rules == W->rules && R->rules == W->rules, "Physical Indices don't match!");
  rules = W->rules;
  // the contractions the reference performs (contract13, contract32, and
  // the psi pairs 30,41 / 20,31,42) must be legal
  CAM_REQUIRE(indices_match(W->idx[1], false, L->idx[3], false) || indices_match(L->idx[3], false, W->idx[1], false),
              "heff: MPO left bond does not match the left block");
  CAM_REQUIRE(indices_match(W->idx[3], false, R->idx[2], false),
              "heff: MPO right bond does not match the right block");
  for (int k : {0, 2, 4}) CAM_REQUIRE(L->idx[k].size == 1, "heff: left block dummy index not size 1");
  for (int k : {1, 3, 5}) CAM_REQUIRE(R->idx[k].size == 1, "heff: right block dummy index not size 1");
  CAM_REQUIRE(L->idx[1].same_charges(L->idx[5]), "heff: left block bra/ket bonds differ");
  CAM_REQUIRE(R->idx[0].same_charges(R->idx[4]), "heff: right block bra/ket bonds differ");
  CAM_REQUIRE(W->idx[0].same_charges(W->idx[2]), "heff: MPO physical indices differ");
  sig = W->idx[2];
  a0 = L->idx[5];
  a2 = R->idx[4];
  d = sig.size;
  chiL = a0.size;
  chiR = a2.size;
  w = W->idx[1].size;
  hipStream_t s = ctx->stream;

  // ---- psi blocks ---------------------------------------------------------
  pblocks.clear();
  pindex.clear();
  n = 0;
  for (int s3 = 0; s3 < a2.nsec(); ++s3)
    for (int s2 = 0; s2 < sig.nsec(); ++s2)
      for (int s1 = 0; s1 < sig.nsec(); ++s1)
        for (int s0 = 0; s0 < a0.nsec(); ++s0) {
          const Charge q = rules.add(rules.add(sig.sec_q[s1], a0.sec_q[s0]), sig.sec_q[s2]);
          if (q != a2.sec_q[s3]) continue;
          PB b{s0, s3, s1, s2, a0.sec_size[s0], a2.sec_size[s3], sig.sec_size[s1], sig.sec_size[s2], n, 0};
          b.size = b.m0 * b.m3 * b.n1 * b.n2;
          pindex[{s0, s3, s1, s2}] = static_cast<int>(pblocks.size());
          pblocks.push_back(b);
          n += b.size;
        }

  // ---- operand views ---------------------------------------------------------
  const int64_t kHostCheckLimit = int64_t(1) << 24;  // elements
  std::vector<cplx> hL, hR;
  const bool chkL = L->total <= kHostCheckLimit, chkR = R->total <= kHostCheckLimit;
  if (chkL) hL = download(*L, s);
  if (chkR) hR = download(*R, s);
  Lv = env_views(*L, 1, 3, 5, hL, chkL);   // (b, s0) -> a0' view
  Rv = env_views(*R, 0, 2, 4, hR, chkR);   // (b, s3) -> a2' view
  std::vector<cplx> hW = download(*W, s);
  // W views: key (bl, br, s', s)
  Wv.clear();
  for (const Block& bl : W->blocks) {
    const int sp = bl.sec[0], sbl = bl.sec[1], ss = bl.sec[2], sbr = bl.sec[3];
    const int64_t no = bl.dims[0], nbl = bl.dims[1], ni = bl.dims[2], nbr = bl.dims[3];
    for (int64_t jl = 0; jl < nbl; ++jl)
      for (int64_t jr = 0; jr < nbr; ++jr) {
        const int64_t base = bl.offset + no * jl + no * nbl * ni * jr;
        bool nz = false;
        for (int64_t k = 0; k < ni && !nz; ++k)
          for (int64_t i = 0; i < no; ++i) {
            const cplx v = hW[base + i + k * no * nbl];
            if (v.re != 0.0 || v.im != 0.0) { nz = true; break; }
          }
        if (!nz) continue;
        const int64_t bl_pos = W->idx[1].sec_pos[sbl][jl];
        const int64_t br_pos = W->idx[3].sec_pos[sbr][jr];
        Wv[{bl_pos, br_pos, sp, ss}] = WMat{base, no * nbl, no, ni};
      }
  }

  // ---- intermediates and the block products of the four stages -----------
  t1.clear(); t2.clear(); t3.clear();
  segA.clear(); segB.clear(); segC.clear(); segD.clear();
  t1n = t2n = t3n = 0;
  for (size_t pi = 0; pi < pblocks.size(); ++pi) {
    const PB& P = pblocks[pi];
    for (int64_t b = 0; b < w; ++b) {
      auto it = Lv.find({b, P.s0});
      if (it == Lv.end()) continue;
      const EnvMat& e = it->second;
      std::array<int, 5> key{static_cast<int>(b), e.sout, P.s3, P.s1, P.s2};
      TB tb{t1n, e.mout, P.m3, P.n1, P.n2};
      t1[key] = tb;
      t1n += tb.m0 * tb.m3 * tb.n1 * tb.n2;
      segA[key] = SegA{static_cast<int>(pi), e};
    }
  }
  // stage B: T2[b1](s0',s3,s1',s2) <- T1[bL](s0',s3,s1,s2) x W(s1',bL,s1,b1)
  for (auto& kv : t1) {
    const auto& k = kv.first;  // (bL, s0', s3, s1, s2)
    for (auto& wv : Wv) {
      const auto& wk = wv.first;  // (bl, br, s', s)
      if (wk[0] != k[0] || wk[3] != k[3]) continue;
      std::array<int, 5> okey{static_cast<int>(wk[1]), k[1], k[2], static_cast<int>(wk[2]), k[4]};
      segB[okey].push_back(SegW{k, wv.second});
    }
  }
  for (auto& kv : segB) {
    const auto& k = kv.first;
    const TB& src = t1[kv.second.front().src];
    TB tb{t2n, src.m0, src.m3, sig.sec_size[k[3]], src.n2};
    t2[k] = tb;
    t2n += tb.m0 * tb.m3 * tb.n1 * tb.n2;
  }
  // stage C: T3[bR](s0',s3,s1',s2') <- T2[b1](s0',s3,s1',s2) x W(s2',b1,s2,bR)
  for (auto& kv : t2) {
    const auto& k = kv.first;  // (b1, s0', s3, s1', s2)
    for (auto& wv : Wv) {
      const auto& wk = wv.first;
      if (wk[0] != k[0] || wk[3] != k[4]) continue;
      std::array<int, 5> okey{static_cast<int>(wk[1]), k[1], k[2], k[3], static_cast<int>(wk[2])};
      segC[okey].push_back(SegW{k, wv.second});
    }
  }
  for (auto& kv : segC) {
    const auto& k = kv.first;
    const TB& src = t2[kv.second.front().src];
    TB tb{t3n, src.m0, src.m3, src.n1, sig.sec_size[k[4]]};
    t3[k] = tb;
    t3n += tb.m0 * tb.m3 * tb.n1 * tb.n2;
  }
  // stage D: out(s0',s3',s1',s2') <- T3[bR](s0',s3,s1',s2') x R_bR(s3'<-s3)
  for (auto& kv : t3) {
    const auto& k = kv.first;  // (bR, s0', s3, s1', s2')
    auto it = Rv.find({k[0], k[2]});
    if (it == Rv.end()) continue;
    const EnvMat& e = it->second;
    auto pit = pindex.find({k[1], e.sout, k[3], k[4]});
    CAM_REQUIRE(pit != pindex.end(), "heff: output block outside the allowed space (flux)");
    segD[pit->second].push_back(SegR{k, e});
  }
  // output blocks that receive nothing must be zeroed after the matvec
  untouched.clear();
  for (size_t pi = 0; pi < pblocks.size(); ++pi)
    if (!segD.count(static_cast<int>(pi))) untouched.push_back(static_cast<int>(pi));

  if (workspace_limit < 0.0) {
    // free device memory plus the context workspaces (reused, grow-only),
    // less a reserve for the caller's vectors and library scratch
    size_t fr = 0, tot = 0;
    HIP_CHECK(hipMemGetInfo(&fr, &tot));
    const double have = double(fr) + sizeof(cplx) * double(ctx->ws_t1.size() + ctx->ws_t2.size());
    workspace_limit = have - 6.0e9;   // the plans' tile lists, library scratch
  }
  choose_slices(workspace_limit);
  plan_slices(rslices, cslices);
  HIP_CHECK(hipStreamSynchronize(s));
}

namespace {
// rows [lo, hi) of slice i out of k of a sector of m rows
inline void slice_range(int64_t m, int i, int k, int64_t& lo, int64_t& cnt) {
  lo = m * i / k;
  cnt = m * (i + 1) / k - lo;
}
}  // namespace

double Heff::slice_workspace_bytes(int rs, int cs) const {
  // per slice, every intermediate block keeps its (a0' rows, a2 columns)
  // share; the buffers must hold the largest slice
  int64_t b1 = 0, b2 = 0;
  for (int ri = 0; ri < rs; ++ri)
    for (int ci = 0; ci < cs; ++ci) {
      auto part = [&](const std::map<std::array<int, 5>, TB>& T) {
        int64_t tot = 0;
        for (auto& kv : T) {
          int64_t r0, mr, c0, mc;
          slice_range(kv.second.m0, ri, rs, r0, mr);
          slice_range(kv.second.m3, ci, cs, c0, mc);
          tot += mr * mc * kv.second.n1 * kv.second.n2;
        }
        return tot;
      };
      b1 = std::max(b1, std::max(part(t1), part(t3)));
      b2 = std::max(b2, part(t2));
    }
  return double(sizeof(cplx)) * double(b1 + b2);
}

void Heff::choose_slices(double bytes) {
  // fewest slices that fit; among equal counts the least extra traffic
  // (psi read once per row slice, the output read and written once per
  // column slice after the first)
  int br = 1, bc = 1;
  bool found = false;
  for (int count = 1; count <= 4096 && !found; count *= 2) {
    double best = 1e300;
    for (int rs = 1; rs <= count; rs *= 2) {
      const int cs = count / rs;
      if (slice_workspace_bytes(rs, cs) > bytes) continue;
      const double traffic = double(rs) + 2.0 * double(cs - 1);
      if (traffic < best) {
        best = traffic;
        br = rs;
        bc = cs;
        found = true;
      }
    }
  }
  CAM_REQUIRE(found, "heff: the intermediates do not fit the workspace limit even in 4096 slices");
  rslices = br;
  cslices = bc;
}

void Heff::plan_slices(int rs, int cs) {
  CAM_REQUIRE(rs >= 1 && cs >= 1, "heff: slice counts must be >= 1");
  rslices = rs;
  cslices = cs;
  plans.clear();
  plans.resize(static_cast<size_t>(rs) * cs);
  ws1n = ws2n = 0;
  flops = 0;
  for (int i = 0; i < 4; ++i) stage_flops[i] = 0;
  auto off = [](int64_t o) { return reinterpret_cast<const void*>(o); };
  auto offw = [](int64_t o) { return reinterpret_cast<void*>(o); };
  for (int ri = 0; ri < rs; ++ri)
    for (int ci = 0; ci < cs; ++ci) {
      std::array<GemmPlan, 4>& st = plans[static_cast<size_t>(ri) * cs + ci];
      // slice-local layouts of the intermediates: block (b, s0', s3, sx, sy)
      // keeps rows [r0, r0+mr) of its a0' sector and columns [c0, c0+mc) of
      // its a2 sector, dense over (a0', a2, sx, sy)
      struct SB { int64_t off, r0, mr, c0, mc, n1, n2; };
      // (blocks in the order of the whole layout, so that 1 x 1 reproduces
      // it exactly: the adjoint reads T1 and T2 at those offsets)
      auto layout = [&](const std::map<std::array<int, 5>, TB>& T, std::map<std::array<int, 5>, SB>& out) {
        std::vector<std::pair<int64_t, const std::array<int, 5>*>> order;
        order.reserve(T.size());
        for (auto& kv : T) order.emplace_back(kv.second.off, &kv.first);
        std::sort(order.begin(), order.end());
        int64_t tot = 0;
        for (auto& o : order) {
          const TB& tb = T.at(*o.second);
          SB b{tot, 0, 0, 0, 0, tb.n1, tb.n2};
          slice_range(tb.m0, ri, rs, b.r0, b.mr);
          slice_range(tb.m3, ci, cs, b.c0, b.mc);
          out[*o.second] = b;
          tot += b.mr * b.mc * b.n1 * b.n2;
        }
        return tot;
      };
      std::map<std::array<int, 5>, SB> l1, l2, l3;
      const int64_t n1 = layout(t1, l1), n2 = layout(t2, l2), n3 = layout(t3, l3);
      ws1n = std::max(ws1n, std::max(n1, n3));
      ws2n = std::max(ws2n, n2);
      // stage A: T1 rows of the slice from L_b's rows, psi's columns of the slice
      for (auto& kv : segA) {
        const SB& tb = l1.at(kv.first);
        if (!tb.mr || !tb.mc) continue;
        const PB& Pb = pblocks[kv.second.pb];
        const EnvMat& e = kv.second.L;
        if (tb.mc == Pb.m3) {
          const int N = static_cast<int>(Pb.m3 * Pb.n1 * Pb.n2);
          const int pi = st[0].add_problem(offw(tb.off), tb.mr, static_cast<int>(tb.mr), N, 1, 0, 0, kSpT1);
          st[0].add_segment(pi, off(e.off + tb.r0), e.ld, 0, off(Pb.off), Pb.m0, 0, static_cast<int>(Pb.m0),
                            seg_spaces(kSpL, kSpIn));
        } else {
          // a2 columns of one (s1, s2) slab are contiguous: batch over the slabs
          const int pi = st[0].add_problem(offw(tb.off), tb.mr, static_cast<int>(tb.mr), static_cast<int>(tb.mc),
                                           static_cast<int>(Pb.n1 * Pb.n2), tb.mr * tb.mc, 0, kSpT1);
          st[0].add_segment(pi, off(e.off + tb.r0), e.ld, 0, off(Pb.off + Pb.m0 * tb.c0), Pb.m0, Pb.m0 * Pb.m3,
                            static_cast<int>(Pb.m0), seg_spaces(kSpL, kSpIn));
        }
      }
      // W stages: the outputs of every MPO channel and new sector with one
      // row space read the same operand rows (row group)
      std::map<std::array<int, 3>, int64_t> rgB, rgC;
      auto row_group = [](std::map<std::array<int, 3>, int64_t>& m, std::array<int, 3> k) {
        return m.emplace(k, static_cast<int64_t>(m.size())).first->second;
      };
      for (auto& kv : segB) {
        const SB& tb = l2.at(kv.first);
        const int64_t rows = tb.mr * tb.mc;
        if (!rows) continue;
        const int pi = st[1].add_problem(offw(tb.off), rows, static_cast<int>(rows), static_cast<int>(tb.n1),
                                         static_cast<int>(tb.n2), rows * tb.n1, 0, kSpT2);
        st[1].set_group(pi, row_group(rgB, {kv.first[1], kv.first[2], kv.first[4]}));
        for (const SegW& sg : kv.second) {
          const SB& sb = l1.at(sg.src);
          st[1].add_segment(pi, off(sb.off), rows, rows * sb.n1, off(sg.W.off), sg.W.ld, 0,
                            static_cast<int>(sb.n1), kSegBNContig | seg_spaces(kSpT1, kSpW));
        }
      }
      for (auto& kv : segC) {
        const SB& tb = l3.at(kv.first);
        const int64_t rows = tb.mr * tb.mc * tb.n1;
        if (!rows) continue;
        const int pi = st[2].add_problem(offw(tb.off), rows, static_cast<int>(rows), static_cast<int>(tb.n2), 1,
                                         0, 0, kSpT3);
        st[2].set_group(pi, row_group(rgC, {kv.first[1], kv.first[2], kv.first[3]}));
        for (const SegW& sg : kv.second) {
          const SB& sb = l2.at(sg.src);
          st[2].add_segment(pi, off(sb.off), rows, 0, off(sg.W.off), sg.W.ld, 0, static_cast<int>(sb.n2),
                            kSegBNContig | seg_spaces(kSpT2, kSpW));
        }
      }
      // stage D: rows of the slice of every output block; the first column
      // slice writes, the others accumulate their part of the a2 sum
      for (auto& kv : segD) {
        const PB& Ob = pblocks[kv.first];
        int64_t r0, mr;
        slice_range(Ob.m0, ri, rs, r0, mr);
        if (!mr) continue;
        const int pi = st[3].add_problem(offw(Ob.off + r0), Ob.m0, static_cast<int>(mr), static_cast<int>(Ob.m3),
                                         static_cast<int>(Ob.n1 * Ob.n2), Ob.m0 * Ob.m3, ci > 0 ? 1 : 0, kSpOut);
        for (const SegR& sg : kv.second) {
          const SB& sb = l3.at(sg.src);
          if (!sb.mc) continue;
          st[3].add_segment(pi, off(sb.off), sb.mr, sb.mr * sb.mc, off(sg.R.off + sg.R.ld * sb.c0), sg.R.ld, 0,
                            static_cast<int>(sb.mc), kSegBNContig | seg_spaces(kSpT3, kSpR));
        }
      }
      for (int i = 0; i < 4; ++i) {
        st[i].choose_tiles(1);
        st[i].build_tiles();
        st[i].upload(ctx->stream);
        stage_flops[i] += st[i].flops;
        flops += st[i].flops;
      }
    }
}

// The intermediates of an FP32 matvec take 8 bytes per element, half the
// FP64 size; the context workspaces are grow-only, so they are sized at the
// first use of a precision (an FP32-only caller never holds the FP64 size).
void Heff::bind_workspace(int precision) {
  const int64_t e1 = precision ? (ws1n + 1) / 2 : ws1n, e2 = precision ? (ws2n + 1) / 2 : ws2n;
  ctx->ws_t1.ensure(std::max<int64_t>(e1, 1));
  ctx->ws_t2.ensure(std::max<int64_t>(e2, 1));
  T1.p = ctx->ws_t1.get();
  T2.p = ctx->ws_t2.get();
}

GemmBases Heff::bases(int precision, const void* in, void* out) {
  // another H_eff on this context may have grown the shared workspaces
  bind_workspace(precision);
  GemmBases b{};
  if (precision == 0) {
    b.p[kSpL] = L->data.get();
    b.p[kSpW] = W->data.get();
    b.p[kSpR] = R->data.get();
    b.p[kSpT1] = T1.get();
    b.p[kSpT2] = T2.get();
    b.p[kSpT3] = T1.get();
  } else {
    b.p[kSpL] = L32.get();
    b.p[kSpW] = W32.get();
    b.p[kSpR] = R32.get();
    b.p[kSpT1] = T1.get();   // read as float2: offsets count FP32 elements
    b.p[kSpT2] = T2.get();
    b.p[kSpT3] = T1.get();
  }
  b.p[kSpIn] = in;
  b.p[kSpOut] = out;
  return b;
}

void Heff::prepare_fp32() {
  if (fp32_ready) return;
  hipStream_t s = ctx->stream;
  L32.alloc(std::max<int64_t>(L->total, 1));
  W32.alloc(std::max<int64_t>(W->total, 1));
  R32.alloc(std::max<int64_t>(R->total, 1));
  launch_to_fp32(L->data.get(), L32.get(), L->total, s);
  launch_to_fp32(W->data.get(), W32.get(), W->total, s);
  launch_to_fp32(R->data.get(), R32.get(), R->total, s);
  HIP_CHECK(hipStreamSynchronize(s));
  fp32_ready = true;
}

namespace {
bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}
}  // namespace

// The first matvec at a precision (outside graph capture) autotunes every
// stage plan of the first slice on its real operands: stage i is re-run
// while stage i-1's output is already in place, and the last run leaves
// the correct result (the first slice's stage D writes, it never
// accumulates).  The other slices have the same shapes up to a row or
// column and take the first slice's configuration.
void Heff::run_stages(int precision, const GemmBases& b) {
  hipStream_t s = ctx->stream;
  const bool tune = !tuned(precision) && !stream_capturing(s);
  for (size_t k = 0; k < plans.size(); ++k) {
    std::array<GemmPlan, 4>& st = plans[k];
    for (int i = 0; i < 4; ++i) {
      if (tune && k == 0) autotune(st[i], precision, s, &b);
      if (tune && k > 0 && plans[0][i].tuned[precision]) {
        st[i].cfg[precision] = plans[0][i].cfg[precision];
        st[i].tuned[precision] = true;
      }
      launch_grouped_gemm(st[i], precision, s, &b);
    }
  }
}

bool Heff::tuned(int precision) const {
  for (const auto& st : plans)
    for (int i = 0; i < 4; ++i)
      if (!st[i].tuned[precision] && plans[0][i].flops >= 2e10) return false;
  return true;
}

void Heff::set_launch(int precision, const LaunchCfg& c) {
  for (auto& st : plans)
    for (int i = 0; i < 4; ++i) {
      st[i].cfg[precision] = c;
      st[i].tuned[precision] = true;
      if (!st[i].built[c.tile_mode]) st[i].build_tiles_mode(c.tile_mode);
      st[i].upload(ctx->stream);
    }
}

void Heff::apply_blocked(const cplx* in, cplx* out) {
  hipStream_t s = ctx->stream;
  run_stages(0, bases(0, in, out));
  for (int pi : untouched) launch_zero(out + pblocks[pi].off, pblocks[pi].size * sizeof(cplx), s);
  if (nproj) add_projectors(in, out);
}

void Heff::set_projectors(int np, const cplx* const* P_ref, const double* weights) {
  CAM_REQUIRE(np >= 0 && np <= 64, "at most 64 projector tensors");
  hipStream_t s = ctx->stream;
  nproj = 0;
  if (np == 0) return;
  build_perm();
  projQ.alloc(static_cast<size_t>(n) * np);
  DevBuf<cplx> tmp(n);
  std::vector<cplx> q(n), w(np);
  for (int p = 0; p < np; ++p) {
    for (int64_t i = 0; i < n; ++i) q[i] = cplx{P_ref[p][i].re, -P_ref[p][i].im};   // Q = conj(P)
    tmp.upload(q.data(), q.size(), s);
    launch_scatter_map(tmp.get(), projQ.get() + static_cast<int64_t>(p) * n, perm.get(), n, s);
    w[p] = cplx{-weights[p], 0.0};
    HIP_CHECK(hipStreamSynchronize(s));
  }
  projW.alloc(np);
  projW.upload(w.data(), w.size(), s);
  projC.alloc(64);
  projPart.alloc(static_cast<size_t>(multi_dot_blocks(n)) * 64);
  projQ32.alloc(static_cast<size_t>(n) * np);
  launch_to_fp32(projQ.get(), projQ32.get(), n * np, s);
  HIP_CHECK(hipStreamSynchronize(s));
  nproj = np;
}

void Heff::add_projectors(const cplx* in, cplx* out) {
  hipStream_t s = ctx->stream;
  launch_multi_dot(projQ.get(), n, nproj, in, n, projPart.get(), projC.get(), s);
  launch_coef_scale(projC.get(), projW.get(), nproj, s);
  launch_multi_axpy(projQ.get(), n, nproj, projC.get(), out, n, s);
}

void Heff::add_projectors32(const float2* in, float2* out) {
  hipStream_t s = ctx->stream;
  launch_multi_dot32(projQ32.get(), n, nproj, in, n, projPart.get(), projC.get(), s);
  launch_coef_scale(projC.get(), projW.get(), nproj, s);
  launch_multi_axpy32(projQ32.get(), n, nproj, projC.get(), out, n, s);
}

void Heff::forward_stages(const cplx* in, int nstages) {
  CAM_REQUIRE(nslices() == 1, "heff: the adjoint needs the whole intermediates (an unsliced matvec)");
  CAM_REQUIRE(ws1n == std::max(t1n, t3n) && ws2n == t2n, "heff: unsliced plan layout differs from the intermediates");
  const GemmBases b = bases(0, in, nullptr);
  for (int i = 0; i < std::min(nstages, 3); ++i) launch_grouped_gemm(plans[0][i], 0, ctx->stream, &b);
}

void Heff::apply_blocked32(const float2* in, float2* out) {
  prepare_fp32();
  hipStream_t s = ctx->stream;
  run_stages(1, bases(1, in, out));
  for (int pi : untouched) launch_zero(out + pblocks[pi].off, pblocks[pi].size * sizeof(float2), s);
  if (nproj) add_projectors32(in, out);
}

void Heff::apply(const cplx* in, cplx* out, int precision) {
  if (precision == 0) {
    apply_blocked(in, out);
    return;
  }
  prepare_fp32();
  hipStream_t s = ctx->stream;
  vinf.ensure(std::max<int64_t>(n, 1));
  voutf.ensure(std::max<int64_t>(n, 1));
  launch_to_fp32(in, vinf.get(), n, s);
  apply_blocked32(vinf.get(), voutf.get());
  launch_to_fp64(voutf.get(), out, n, s);
}

TensorPtr Heff::tensor_from_vector(const cplx* x) {
  auto t = std::make_shared<BSTensor>();
  t->rules = rules;
  QIndex s1 = sig, s2 = sig, i0 = a0, i2 = a2;
  s1.inward = s2.inward = true;
  i0.inward = true;
  i2.inward = false;
  t->idx = {s1, i0, s2, i2};
  t->numrows = 2;
  for (const PB& P : pblocks) {
    Block b;
    b.sec = {P.s1, P.s0, P.s2, P.s3};
    b.dims = {P.n1, P.m0, P.n2, P.m3};
    t->blocks.push_back(b);
  }
  std::sort(t->blocks.begin(), t->blocks.end(),
            [&](const Block& a, const Block& b) { return t->code(a.sec) < t->code(b.sec); });
  t->finalize_layout();
  std::vector<PermJob> jobs;
  for (const PB& P : pblocks) {
    const Block& o = t->blocks[t->find_block({P.s1, P.s0, P.s2, P.s3})];
    PermJob j{};
    j.in = x + P.off;
    j.out = t->data.get() + o.offset;
    j.size = P.size;
    j.ndim = 4;
    // output (s1, a0, s2, a2) <- block layout (a0, a2, s1, s2)
    j.odims[0] = P.n1; j.istr[0] = P.m0 * P.m3;
    j.odims[1] = P.m0; j.istr[1] = 1;
    j.odims[2] = P.n2; j.istr[2] = P.m0 * P.m3 * P.n1;
    j.odims[3] = P.m3; j.istr[3] = P.m0;
    if (j.size) jobs.push_back(j);
  }
  launch_permute(jobs, ctx->perm_jobs, ctx->perm_chunks, ctx->stream);
  return t;
}

void Heff::vector_from_tensor(const BSTensor& T, cplx* x) {
  CAM_REQUIRE(T.nidx() == 4, "wavefunction must have 4 indices");
  CAM_REQUIRE(T.idx[0].same_charges(sig) && T.idx[2].same_charges(sig) && T.idx[1].same_charges(a0) &&
                  T.idx[3].same_charges(a2),
              "wavefunction indices do not match the effective Hamiltonian");
  launch_zero(x, n * sizeof(cplx), ctx->stream);
  std::vector<PermJob> jobs;
  for (const Block& b : T.blocks) {
    auto it = pindex.find({b.sec[1], b.sec[3], b.sec[0], b.sec[2]});
    CAM_REQUIRE(it != pindex.end(), "wavefunction block outside the allowed space");
    const PB& P = pblocks[it->second];
    PermJob j{};
    j.in = T.data.get() + b.offset;
    j.out = x + P.off;
    j.size = P.size;
    j.ndim = 4;
    // output (a0, a2, s1, s2) <- tensor layout (s1, a0, s2, a2)
    j.odims[0] = P.m0; j.istr[0] = P.n1;
    j.odims[1] = P.m3; j.istr[1] = P.n1 * P.m0 * P.n2;
    j.odims[2] = P.n1; j.istr[2] = 1;
    j.odims[3] = P.n2; j.istr[3] = P.n1 * P.m0;
    if (j.size) jobs.push_back(j);
  }
  launch_permute(jobs, ctx->perm_jobs, ctx->perm_chunks, ctx->stream);
}

void Heff::profile(int precision, int reps, double* stage_ms, const void* in_dev, void* out_dev) {
  hipStream_t s = ctx->stream;
  if (precision == 1) prepare_fp32();
  DevBuf<cplx> x, y;
  DevBuf<float2> xf, yf;
  const void* in = in_dev;
  void* out = out_dev;
  if (!in || !out) {
    x.alloc(std::max<int64_t>(n, 1));
    y.alloc(std::max<int64_t>(n, 1));
    // deterministic non-zero input
    std::vector<cplx> h(static_cast<size_t>(std::min<int64_t>(n, 1 << 20)));
    for (size_t i = 0; i < h.size(); ++i) h[i] = cplx{std::sin(0.37 * i), std::cos(0.11 * i)};
    for (int64_t o = 0; o < n; o += static_cast<int64_t>(h.size()))
      HIP_CHECK(hipMemcpyAsync(x.get() + o, h.data(), std::min<int64_t>(h.size(), n - o) * sizeof(cplx),
                               hipMemcpyHostToDevice, s));
    in = x.get();
    out = y.get();
    if (precision == 1) {
      xf.alloc(std::max<int64_t>(n, 1));
      yf.alloc(std::max<int64_t>(n, 1));
      launch_to_fp32(x.get(), xf.get(), n, s);
      in = xf.get();
      out = yf.get();
    }
  }
  const GemmBases b = bases(precision, in, out);
  // one event before the matvec and one after every (slice, stage) launch;
  // a stage's time is the sum over the slices of its launches
  const size_t nl = 4 * plans.size();
  std::vector<hipEvent_t> ev(nl + 1);
  for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
  double acc[4] = {0, 0, 0, 0};
  for (int r = 0; r < reps + 1; ++r) {            // first pass is a warmup
    HIP_CHECK(hipEventRecord(ev[0], s));
    for (size_t k = 0; k < plans.size(); ++k)
      for (int i = 0; i < 4; ++i) {
        launch_grouped_gemm(plans[k][i], precision, s, &b);
        HIP_CHECK(hipEventRecord(ev[4 * k + i + 1], s));
      }
    HIP_CHECK(hipEventSynchronize(ev[nl]));
    if (r == 0) continue;
    for (size_t j = 0; j < nl; ++j) {
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, ev[j], ev[j + 1]));
      acc[j % 4] += ms;
    }
  }
  for (auto& e : ev) HIP_CHECK(hipEventDestroy(e));
  for (int i = 0; i < 4; ++i) stage_ms[i] = reps > 0 ? acc[i] / reps : 0.0;
}

void Heff::build_perm() {
  if (perm_ready) return;
  // reference compressed order (DMRG_routines.cpp:808-824): loop a2 > s2 >
  // a0 > s1, keep entries with sigma1 == a2 - sigma2 - a0; flat index
  // s1 + d*a0 + d*chiL*(s2 + d*a2) is ascending in that loop order.
  allowed.clear();
  allowed.reserve(static_cast<size_t>(n));
  std::vector<int64_t> pm;
  pm.reserve(static_cast<size_t>(n));
  for (int64_t ia2 = 0; ia2 < chiR; ++ia2) {
    const int s3 = a2.pos_sec[ia2];
    const int64_t o3 = a2.pos_off[ia2];
    for (int64_t is2 = 0; is2 < d; ++is2) {
      const int s2 = sig.pos_sec[is2];
      const int64_t o2 = sig.pos_off[is2];
      const Charge acc2 = rules.sub(a2.charges[ia2], sig.charges[is2]);
      for (int64_t ia0 = 0; ia0 < chiL; ++ia0) {
        const Charge need = rules.sub(acc2, a0.charges[ia0]);
        const int s1 = sig.find_sector(need);
        if (s1 < 0) continue;
        const int s0 = a0.pos_sec[ia0];
        const int64_t o0 = a0.pos_off[ia0];
        auto it = pindex.find({s0, s3, s1, s2});
        CAM_REQUIRE(it != pindex.end(), "heff: allowed entry without a block");
        const PB& P = pblocks[it->second];
        const int64_t colbase = d * chiL * (is2 + d * ia2) + d * ia0;
        for (int64_t o1 = 0; o1 < P.n1; ++o1) {
          const int64_t is1 = sig.sec_pos[s1][o1];
          allowed.push_back(is1 + colbase);
          pm.push_back(P.off + o0 + P.m0 * (o3 + P.m3 * (o1 + P.n1 * o2)));
        }
      }
    }
  }
  CAM_REQUIRE(static_cast<int64_t>(allowed.size()) == n, "heff: allowed count mismatch");
  // rows with the same (a2,s2,a0) were emitted in sector order == ascending
  // position order, so `allowed` is ascending
  perm.alloc(std::max<int64_t>(n, 1));
  perm.upload(pm.data(), pm.size(), ctx->stream);
  HIP_CHECK(hipStreamSynchronize(ctx->stream));
  perm_ready = true;
}

}  // namespace cam
