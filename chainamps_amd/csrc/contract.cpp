// ajaj::contract / reorder on QN-blocked device tensors.
//
// contract (MPX_matrix.cpp:640-791): result indices are A's kept indices in
// order (flipped if conjA) then B's kept indices (flipped if conjB); the
// contracted indices are matched in pair order.  Every A block is paired
// with every B block whose contracted sectors agree; all pairs feeding the
// same output block become K-segments of one problem of a single grouped
// MFMA launch.  Operands already laid out as (kept, contracted) for A and
// (contracted, kept) or (kept, contracted) for B are used in place;
// otherwise blocks are permuted once into scratch.
#include <algorithm>
#include <map>
#include <numeric>

#include "ctx.hpp"

namespace cam {

Ctx::Ctx(int dev) : device(dev) {
  HIP_CHECK(hipSetDevice(dev));
  HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
}

bool indices_match(const QIndex& a, bool conjA, const QIndex& b, bool conjB) {
  // MPXIndex.cpp:47-72 (charges compared always, the reference does it in
  // debug builds)
  if (a.physical != b.physical) return false;
  const bool aout = !a.inward, bin = b.inward;
  if ((aout != conjA) != (bin != conjB)) return false;
  if (a.size != b.size) return false;
  return a.charges == b.charges;
}

void permute_tensor(Ctx& ctx, const BSTensor& A, const std::vector<int>& order, bool conj,
                    const BSTensor& out_layout, cplx* out) {
  const int n = A.nidx();
  CAM_REQUIRE(n <= kMaxDims, "too many indices for a permutation");
  std::vector<PermJob> jobs;
  std::vector<int32_t> sec(n);
  for (const Block& b : A.blocks) {
    for (int k = 0; k < n; ++k) sec[k] = b.sec[order[k]];
    const int ob = out_layout.find_block(sec);
    CAM_REQUIRE(ob >= 0, "permute: output block missing");
    const Block& o = out_layout.blocks[ob];
    PermJob j{};
    j.in = A.data.get() + b.offset;
    j.out = out + o.offset;
    j.size = b.size;
    j.ndim = n;
    j.conj = conj ? 1 : 0;
    std::vector<int64_t> istr(n);
    int64_t st = 1;
    for (int k = 0; k < n; ++k) { istr[k] = st; st *= b.dims[k]; }
    for (int k = 0; k < n; ++k) {
      j.odims[k] = o.dims[k];
      j.istr[k] = istr[order[k]];
    }
    if (j.size) jobs.push_back(j);
  }
  launch_permute(jobs, ctx.perm_jobs, ctx.perm_chunks, ctx.stream);
}

// Layout of A's blocks under an index permutation (same block set).
static TensorPtr permuted_layout(const BSTensor& A, const std::vector<int>& order, bool flip,
                                 int numrows) {
  auto t = std::make_shared<BSTensor>();
  t->rules = A.rules;
  for (int k : order) t->idx.push_back(flip ? A.idx[k].flipped() : A.idx[k]);
  t->numrows = numrows;
  for (const Block& b : A.blocks) {
    Block nb;
    for (int k : order) { nb.sec.push_back(b.sec[k]); nb.dims.push_back(b.dims[k]); }
    t->blocks.push_back(std::move(nb));
  }
  std::sort(t->blocks.begin(), t->blocks.end(), [&](const Block& x, const Block& y) {
    return t->code(x.sec) < t->code(y.sec);
  });
  return t;
}

TensorPtr reorder(Ctx& ctx, const BSTensor& A, bool conjA, const std::vector<int>& order,
                  int numrows) {
  CAM_REQUIRE(static_cast<int>(order.size()) == A.nidx(), "Incorrect MPX reshape params");
  std::vector<int> seen(order.size(), 0);
  for (int o : order) {
    CAM_REQUIRE(o >= 0 && o < A.nidx() && !seen[o], "reorder: not a permutation");
    seen[o] = 1;
  }
  CAM_REQUIRE(numrows >= 0 && numrows <= A.nidx(), "reorder: bad num_row_indices");
  TensorPtr t = permuted_layout(A, order, conjA, numrows);
  t->finalize_layout();
  permute_tensor(ctx, A, order, conjA, *t, t->data.get());
  return t;
}

TensorPtr contract(Ctx& ctx, const BSTensor& A, bool conjA, const BSTensor& B, bool conjB,
                   const std::vector<std::pair<int, int>>& pairs) {
  CAM_REQUIRE(A.rules == B.rules, "Physical Indices don't match!");
  const int na = A.nidx(), nb = B.nidx();
  std::vector<int> ca, cb;
  {
    std::vector<int> ua(na, 0), ub(nb, 0);
    for (auto& pr : pairs) {
      CAM_REQUIRE(pr.first >= 0 && pr.first < na && pr.second >= 0 && pr.second < nb,
                  "contraction index out of range");
      CAM_REQUIRE(!ua[pr.first] && !ub[pr.second], "Illegal contraction, repeated indices!");
      ua[pr.first] = ub[pr.second] = 1;
      CAM_REQUIRE(indices_match(A.idx[pr.first], conjA, B.idx[pr.second], conjB),
                  "Indices don't match!");
      ca.push_back(pr.first);
      cb.push_back(pr.second);
    }
  }
  std::vector<int> ka, kb;
  for (int i = 0; i < na; ++i)
    if (std::find(ca.begin(), ca.end(), i) == ca.end()) ka.push_back(i);
  for (int i = 0; i < nb; ++i)
    if (std::find(cb.begin(), cb.end(), i) == cb.end()) kb.push_back(i);

  auto out = std::make_shared<BSTensor>();
  out->rules = A.rules;
  for (int i : ka) out->idx.push_back(conjA ? A.idx[i].flipped() : A.idx[i]);
  for (int i : kb) out->idx.push_back(conjB ? B.idx[i].flipped() : B.idx[i]);
  out->numrows = static_cast<int>(ka.size());
  CAM_REQUIRE(out->nidx() <= 16, "too many indices");

  // operand layouts
  std::vector<int> orderA = ka;
  orderA.insert(orderA.end(), ca.begin(), ca.end());
  std::vector<int> orderBk = cb;
  orderBk.insert(orderBk.end(), kb.begin(), kb.end());
  std::vector<int> orderBn = kb;
  orderBn.insert(orderBn.end(), cb.begin(), cb.end());
  auto is_identity = [](const std::vector<int>& o) {
    for (size_t i = 0; i < o.size(); ++i)
      if (o[i] != static_cast<int>(i)) return false;
    return true;
  };
  const bool readyA = is_identity(orderA);
  const bool readyBk = is_identity(orderBk);
  const bool readyBn = !readyBk && is_identity(orderBn);

  // group B blocks by contracted sectors (pair order)
  std::map<std::vector<int32_t>, std::vector<int>> bgroups;
  for (size_t b = 0; b < B.blocks.size(); ++b) {
    std::vector<int32_t> key;
    for (int i : cb) key.push_back(B.blocks[b].sec[i]);
    bgroups[key].push_back(static_cast<int>(b));
  }
  struct Seg { int a, b; };
  std::map<std::vector<int32_t>, std::vector<Seg>> outsegs;
  for (size_t a = 0; a < A.blocks.size(); ++a) {
    std::vector<int32_t> key;
    for (int i : ca) key.push_back(A.blocks[a].sec[i]);
    auto it = bgroups.find(key);
    if (it == bgroups.end()) continue;
    for (int b : it->second) {
      std::vector<int32_t> osec;
      for (int i : ka) osec.push_back(A.blocks[a].sec[i]);
      for (int i : kb) osec.push_back(B.blocks[b].sec[i]);
      outsegs[osec].push_back(Seg{static_cast<int>(a), b});
    }
  }
  for (auto& kv : outsegs) {
    Block ob;
    ob.sec = kv.first;
    for (int k = 0; k < out->nidx(); ++k) ob.dims.push_back(out->idx[k].sec_size[ob.sec[k]]);
    out->blocks.push_back(std::move(ob));
  }
  std::sort(out->blocks.begin(), out->blocks.end(), [&](const Block& x, const Block& y) {
    return out->code(x.sec) < out->code(y.sec);
  });
  out->finalize_layout();
  if (out->total == 0) return out;

  // permuted operands
  const cplx* Abase = A.data.get();
  TensorPtr Alay;
  if (!readyA) {
    Alay = permuted_layout(A, orderA, false, 0);
    int64_t off = 0;
    for (auto& b : Alay->blocks) {
      b.size = 1;
      for (auto d : b.dims) b.size *= d;
      b.offset = off;
      off += b.size;
    }
    Alay->total = off;
    Alay->rebuild_lookup();
    ctx.scratchA.ensure(static_cast<size_t>(std::max<int64_t>(off, 1)));
    permute_tensor(ctx, A, orderA, false, *Alay, ctx.scratchA.get());
    Abase = ctx.scratchA.get();
  }
  const cplx* Bbase = B.data.get();
  TensorPtr Blay;
  bool bncontig = readyBn;
  if (!readyBk && !readyBn) {
    Blay = permuted_layout(B, orderBk, false, 0);
    int64_t off = 0;
    for (auto& b : Blay->blocks) {
      b.size = 1;
      for (auto d : b.dims) b.size *= d;
      b.offset = off;
      off += b.size;
    }
    Blay->total = off;
    Blay->rebuild_lookup();
    ctx.scratchB.ensure(static_cast<size_t>(std::max<int64_t>(off, 1)));
    permute_tensor(ctx, B, orderBk, false, *Blay, ctx.scratchB.get());
    Bbase = ctx.scratchB.get();
  }

  auto prod = [](const std::vector<int64_t>& d, const std::vector<int>& which,
                 const std::vector<int32_t>&) {
    int64_t p = 1;
    for (int i : which) p *= d[i];
    return p;
  };
  GemmPlan plan;
  const int flags = (conjA ? kSegConjA : 0) | (conjB ? kSegConjB : 0) | (bncontig ? kSegBNContig : 0);
  for (auto& kv : outsegs) {
    const int obn = out->find_block(kv.first);
    const Block& ob = out->blocks[obn];
    const Block& a0 = A.blocks[kv.second.front().a];
    const int64_t M = prod(a0.dims, ka, a0.sec);
    int64_t N = 1;
    {
      const Block& b0 = B.blocks[kv.second.front().b];
      N = prod(b0.dims, kb, b0.sec);
    }
    CAM_REQUIRE(M < (int64_t(1) << 31) && N < (int64_t(1) << 31), "block too large");
    const int pi = plan.add_problem(out->data.get() + ob.offset, M, static_cast<int>(M),
                                    static_cast<int>(N), 1, 0, 0);
    for (const Seg& sg : kv.second) {
      const Block& ab = A.blocks[sg.a];
      const Block& bb = B.blocks[sg.b];
      const int64_t K = prod(ab.dims, ca, ab.sec);
      const cplx* Ap;
      if (readyA) {
        Ap = Abase + ab.offset;
      } else {
        std::vector<int32_t> ps;
        for (int i : orderA) ps.push_back(ab.sec[i]);
        Ap = Abase + Alay->blocks[Alay->find_block(ps)].offset;
      }
      const cplx* Bp;
      if (readyBk || readyBn) {
        Bp = Bbase + bb.offset;
      } else {
        std::vector<int32_t> ps;
        for (int i : orderBk) ps.push_back(bb.sec[i]);
        Bp = Bbase + Blay->blocks[Blay->find_block(ps)].offset;
      }
      plan.add_segment(pi, Ap, M, 0, Bp, bncontig ? N : K, 0, static_cast<int>(K), flags);
    }
  }
  // output blocks receiving no segment were not created; zero is implicit
  plan.choose_tiles(1);
  plan.build_tiles();
  launch_grouped_gemm(plan, 0, ctx.stream);
  ctx.contract_flops += plan.flops;
  HIP_CHECK(hipStreamSynchronize(ctx.stream));
  return out;
}

}  // namespace cam
