// Device-resident QN-blocked tensor.
//
// A tensor is an ordered list of QIndex plus the set of present blocks.  A
// block is one sector per index with zero total flux
//   sum_{inward} q - sum_{outward} q == 0,
// which is the reference's zero-flux rule IngoingRowState(r) ==
// OutgoingColState(c) (MPX_matrix.cpp:275-302) rewritten over all indices.
// Each block is a dense array, column-major over the index order (first
// index fastest, as the reference's mixed radix), packed back to back in
// one device buffer.  The row/column split (num_row_indices) is metadata
// only: ShiftNumRowIndices / RemoveDummyIndices / MoveDummyIndices move no
// data in this storage.
#pragma once
#include <cstdint>
#include <memory>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "qn.hpp"

namespace cam {

struct Block {
  std::vector<int32_t> sec;
  std::vector<int64_t> dims;
  int64_t offset = 0;
  int64_t size = 0;
};

struct BSTensor {
  Rules rules;
  std::vector<QIndex> idx;
  int numrows = 0;
  std::vector<Block> blocks;
  std::unordered_map<uint64_t, int> lookup;  // block code -> block number
  DevBuf<cplx> data;
  int64_t total = 0;

  int nidx() const { return static_cast<int>(idx.size()); }
  uint64_t code(const std::vector<int32_t>& sec) const {
    uint64_t c = 0, m = 1;
    for (size_t k = 0; k < sec.size(); ++k) {
      c += static_cast<uint64_t>(sec[k]) * m;
      m *= static_cast<uint64_t>(idx[k].nsec() ? idx[k].nsec() : 1);
    }
    return c;
  }
  int find_block(const std::vector<int32_t>& sec) const {
    auto it = lookup.find(code(sec));
    return it == lookup.end() ? -1 : it->second;
  }
  Charge flux(const std::vector<int32_t>& sec) const {
    Charge acc = rules.zero();
    for (size_t k = 0; k < sec.size(); ++k) {
      const Charge& q = idx[k].sec_q[sec[k]];
      acc = idx[k].inward ? rules.add(acc, q) : rules.sub(acc, q);
    }
    return acc;
  }
  int64_t rows() const {
    int64_t r = 1;
    for (int k = 0; k < numrows; ++k) r *= idx[k].size;
    return r;
  }
  int64_t cols() const {
    int64_t c = 1;
    for (int k = numrows; k < nidx(); ++k) c *= idx[k].size;
    return c;
  }
  // Lay out `blocks` (sec/dims filled) contiguously, build lookup,
  // allocate device memory (uninitialised).
  void finalize_layout();
  void rebuild_lookup();
  // Metadata-only index operations (no data moves in block storage):
  // RemoveDummyIndices (MPX_matrix.cpp:415-457) and MoveDummyIndices
  // (MPX_matrix.cpp:87-105).
  void remove_dummy_indices(const std::vector<int>& which);
  void move_dummy_indices(const std::vector<int>& order);
  // Enumerate every zero-flux sector tuple (used for dense-in-QN outputs).
  std::vector<std::vector<int32_t>> allowed_sector_tuples() const;
};

using TensorPtr = std::shared_ptr<BSTensor>;

// Host-side CSC <-> block conversion.
TensorPtr tensor_from_csc(const Rules& rules, std::vector<QIndex> idx, int numrows, int64_t rows,
                          int64_t cols, const int64_t* p, const int64_t* i, const double* x,
                          hipStream_t s);
// Returns nnz (exact zeros dropped); fills p (cols+1), i, x when non-null
// and cap suffices.
int64_t tensor_to_csc(const BSTensor& t, int64_t cap, int64_t* p, int64_t* i, double* x,
                      hipStream_t s);
std::vector<cplx> download(const BSTensor& t, hipStream_t s);

}  // namespace cam
