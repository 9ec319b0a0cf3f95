#include "bstensor.hpp"

#include <algorithm>
#include <numeric>
#include <functional>

namespace cam {

void BSTensor::rebuild_lookup() {
  lookup.clear();
  for (size_t b = 0; b < blocks.size(); ++b) lookup[code(blocks[b].sec)] = static_cast<int>(b);
}

void BSTensor::finalize_layout() {
  int64_t off = 0;
  for (auto& b : blocks) {
    b.size = 1;
    for (auto d : b.dims) b.size *= d;
    b.offset = off;
    off += b.size;
  }
  total = off;
  rebuild_lookup();
  data.alloc(static_cast<size_t>(total));
}

void BSTensor::remove_dummy_indices(const std::vector<int>& which) {
  CAM_REQUIRE(static_cast<int>(which.size()) < nidx() - 1, "Too many indices for removal!");
  std::vector<int> rm(nidx(), 0);
  for (int w : which) {
    CAM_REQUIRE(w >= 0 && w < nidx() && idx[w].size == 1, "Not a dummy index!");
    rm[w] = 1;
  }
  int nrow_removed = 0;
  for (int k = 0; k < nidx(); ++k)
    if (rm[k] && k < numrows) ++nrow_removed;
  CAM_REQUIRE(nrow_removed < numrows, "All row indices would be removed, requiring a reshape!");
  std::vector<QIndex> ni;
  for (int k = 0; k < nidx(); ++k)
    if (!rm[k]) ni.push_back(idx[k]);
  for (Block& b : blocks) {
    std::vector<int32_t> sc;
    std::vector<int64_t> d;
    for (int k = 0; k < nidx(); ++k)
      if (!rm[k]) { sc.push_back(b.sec[k]); d.push_back(b.dims[k]); }
    b.sec = sc;
    b.dims = d;
  }
  idx = ni;
  numrows -= nrow_removed;
  rebuild_lookup();
}

void BSTensor::move_dummy_indices(const std::vector<int>& order) {
  const int n = nidx();
  CAM_REQUIRE(static_cast<int>(order.size()) == n, "Incorrect MPX reshape params");
  std::vector<int> nondummy;
  std::vector<QIndex> ni;
  int newrows = numrows;
  std::vector<int> seen(n, 0);
  for (int p = 0; p < n; ++p) {
    const int o = order[p];
    CAM_REQUIRE(o >= 0 && o < n && !seen[o], "bad order");
    seen[o] = 1;
    if (idx[o].size > 1) nondummy.push_back(o);
    if (o == numrows) newrows = p;
    ni.push_back(idx[o]);
  }
  for (size_t k = 1; k < nondummy.size(); ++k)
    CAM_REQUIRE(nondummy[k] > nondummy[k - 1], "Non dummy order shouldn't change!");
  for (Block& b : blocks) {
    std::vector<int32_t> sc(n);
    std::vector<int64_t> d(n);
    for (int p = 0; p < n; ++p) { sc[p] = b.sec[order[p]]; d[p] = b.dims[order[p]]; }
    b.sec = sc;
    b.dims = d;
  }
  idx = ni;
  numrows = newrows;
  rebuild_lookup();
}

std::vector<std::vector<int32_t>> BSTensor::allowed_sector_tuples() const {
  std::vector<std::vector<int32_t>> out;
  int n = nidx();
  if (n == 0) return out;
  std::vector<int32_t> cur(n, 0);
  // enumerate sectors of indices 0..n-2, solve for the last one
  std::function<void(int, Charge)> rec = [&](int k, Charge acc) {
    if (k == n - 1) {
      // need: acc + sign*q_last == 0  ->  q_last = inward ? -acc : acc
      Charge need = idx[k].inward ? rules.neg(acc) : acc;
      int s = idx[k].find_sector(need);
      if (s >= 0) {
        cur[k] = s;
        out.push_back(cur);
      }
      return;
    }
    for (int s = 0; s < idx[k].nsec(); ++s) {
      cur[k] = s;
      const Charge& q = idx[k].sec_q[s];
      rec(k + 1, idx[k].inward ? rules.add(acc, q) : rules.sub(acc, q));
    }
  };
  rec(0, rules.zero());
  return out;
}

static void decompose(int64_t flat, const std::vector<int64_t>& dims, int64_t* out) {
  for (size_t k = 0; k < dims.size(); ++k) {
    out[k] = flat % dims[k];
    flat /= dims[k];
  }
}

TensorPtr tensor_from_csc(const Rules& rules, std::vector<QIndex> idx, int numrows, int64_t rows,
                          int64_t cols, const int64_t* p, const int64_t* ii, const double* x,
                          hipStream_t s) {
  auto t = std::make_shared<BSTensor>();
  t->rules = rules;
  t->idx = std::move(idx);
  t->numrows = numrows;
  const int n = t->nidx();
  CAM_REQUIRE(numrows >= 0 && numrows <= n, "bad num_row_indices");
  for (auto& q : t->idx) q.build_sectors();
  CAM_REQUIRE(t->rows() == rows && t->cols() == cols, "CSC shape does not match the indices");
  std::vector<int64_t> rdims, cdims;
  for (int k = 0; k < numrows; ++k) rdims.push_back(t->idx[k].size);
  for (int k = numrows; k < n; ++k) cdims.push_back(t->idx[k].size);
  std::vector<uint64_t> mult(n);
  {
    uint64_t m = 1;
    for (int k = 0; k < n; ++k) { mult[k] = m; m *= static_cast<uint64_t>(t->idx[k].nsec()); }
  }
  // pass 1: discover blocks
  std::unordered_map<uint64_t, int> found;
  std::vector<std::vector<int32_t>> secs;
  std::vector<int64_t> coord(n);
  const int64_t nnz = cols > 0 ? p[cols] : 0;
  std::vector<int32_t> ent_block(static_cast<size_t>(nnz));
  std::vector<int64_t> ent_off(static_cast<size_t>(nnz));
  std::vector<int32_t> sec(n);
  for (int64_t c = 0; c < cols; ++c) {
    if (p[c] == p[c + 1]) continue;
    decompose(c, cdims, coord.data() + numrows);
    for (int64_t e = p[c]; e < p[c + 1]; ++e) {
      int64_t r = ii[e];
      CAM_REQUIRE(r >= 0 && r < rows, "CSC row index out of range");
      decompose(r, rdims, coord.data());
      uint64_t code = 0;
      for (int k = 0; k < n; ++k) {
        sec[k] = t->idx[k].pos_sec[coord[k]];
        code += static_cast<uint64_t>(sec[k]) * mult[k];
      }
      auto it = found.find(code);
      int b;
      if (it == found.end()) {
        b = static_cast<int>(secs.size());
        found.emplace(code, b);
        secs.push_back(sec);
      } else {
        b = it->second;
      }
      ent_block[e] = b;
      // in-block offset (column-major over sector sizes)
      int64_t off = 0, st = 1;
      for (int k = 0; k < n; ++k) {
        off += t->idx[k].pos_off[coord[k]] * st;
        st *= t->idx[k].sec_size[sec[k]];
      }
      ent_off[e] = off;
    }
  }
  // zero flux check (MPX_matrix::isConsistent)
  for (auto& sc : secs) {
    Charge f = t->flux(sc);
    CAM_REQUIRE(f == rules.zero(), "tensor violates the zero-flux rule (isConsistent fails)");
  }
  // deterministic block order: by code
  std::vector<int> order(secs.size());
  std::iota(order.begin(), order.end(), 0);
  std::vector<uint64_t> codes(secs.size());
  for (size_t b = 0; b < secs.size(); ++b) codes[b] = t->code(secs[b]);
  std::sort(order.begin(), order.end(), [&](int a, int b) { return codes[a] < codes[b]; });
  std::vector<int> newno(secs.size());
  for (size_t k = 0; k < order.size(); ++k) {
    newno[order[k]] = static_cast<int>(k);
    Block bl;
    bl.sec = secs[order[k]];
    for (int j = 0; j < n; ++j) bl.dims.push_back(t->idx[j].sec_size[bl.sec[j]]);
    t->blocks.push_back(std::move(bl));
  }
  t->finalize_layout();
  std::vector<cplx> host(static_cast<size_t>(t->total), cplx{0.0, 0.0});
  for (int64_t e = 0; e < nnz; ++e) {
    const Block& bl = t->blocks[newno[ent_block[e]]];
    cplx& dst = host[bl.offset + ent_off[e]];
    dst.re += x[2 * e];
    dst.im += x[2 * e + 1];
  }
  t->data.upload(host.data(), host.size(), s);
  HIP_CHECK(hipStreamSynchronize(s));
  return t;
}

std::vector<cplx> download(const BSTensor& t, hipStream_t s) {
  std::vector<cplx> h(static_cast<size_t>(t.total));
  if (t.total) {
    HIP_CHECK(hipMemcpyAsync(h.data(), t.data.get(), t.total * sizeof(cplx), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
  return h;
}

int64_t tensor_to_csc(const BSTensor& t, int64_t cap, int64_t* p, int64_t* ii, double* x,
                      hipStream_t s) {
  std::vector<cplx> h = download(t, s);
  const int n = t.nidx();
  const int64_t rows = t.rows(), cols = t.cols();
  std::vector<int64_t> rstride(n, 0), cstride(n, 0);
  {
    int64_t m = 1;
    for (int k = 0; k < t.numrows; ++k) { rstride[k] = m; m *= t.idx[k].size; }
    m = 1;
    for (int k = t.numrows; k < n; ++k) { cstride[k] = m; m *= t.idx[k].size; }
  }
  struct Ent { int64_t col, row; cplx v; };
  std::vector<Ent> ents;
  std::vector<int64_t> coord(n);
  for (const Block& bl : t.blocks) {
    for (int64_t e = 0; e < bl.size; ++e) {
      const cplx v = h[bl.offset + e];
      if (v.re == 0.0 && v.im == 0.0) continue;
      int64_t rem = e, r = 0, c = 0;
      for (int k = 0; k < n; ++k) {
        int64_t o = rem % bl.dims[k];
        rem /= bl.dims[k];
        int64_t pos = t.idx[k].sec_pos[bl.sec[k]][o];
        if (k < t.numrows) r += pos * rstride[k]; else c += pos * cstride[k];
      }
      ents.push_back({c, r, v});
    }
  }
  const int64_t nnz = static_cast<int64_t>(ents.size());
  if (!p || nnz > cap) return nnz;
  std::sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) {
    return a.col != b.col ? a.col < b.col : a.row < b.row;
  });
  std::fill(p, p + cols + 1, 0);
  for (auto& e : ents) p[e.col + 1]++;
  for (int64_t c = 0; c < cols; ++c) p[c + 1] += p[c];
  for (int64_t k = 0; k < nnz; ++k) {
    ii[k] = ents[k].row;
    x[2 * k] = ents[k].v.re;
    x[2 * k + 1] = ents[k].v.im;
  }
  (void)rows;
  return nnz;
}

}  // namespace cam
