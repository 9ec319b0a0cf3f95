// Library context: one device, one stream, reusable scratch.
#pragma once
#include <array>
#include <map>
#include <memory>
#include <vector>

#include "bstensor.hpp"
#include "common.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace cam {

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* blas = nullptr;          // rocblas_handle (created lazily)
  DevBuf<cplx> scratchA, scratchB;
  DevBuf<PermJob> perm_jobs;
  DevBuf<int64_t> perm_chunks;
  double contract_flops = 0.0;   // algorithmic 8mnk of every contract() so far
  // Grow-only workspaces reused across calls on this context (one stream,
  // one host thread): the Lanczos basis and the two H_eff intermediates.
  // Freeing and re-allocating ~10^11-byte buffers between eigensolves costs
  // seconds near the memory limit, so they are kept until the context ends.
  DevBuf<cplx> ws_krylov, ws_t1, ws_t2;
  // Leading-values SVD: per block shape (m, n, k) and number of checked
  // triplets, the subspace rounds
  // before the first residual check (the rounds the last accepted block of
  // that shape needed; lowered by one after three blocks in a row were
  // accepted at their first check) and that streak.  It changes only how
  // many rounds run, never the acceptance test.
  std::map<std::array<int64_t, 4>, std::array<int, 2>> svd_round_hint;
  // Leading-values SVD: per block shape (m, n, k) and block charge, how many
  // of its values the last truncation kept (the triplets checked first in
  // the next block of that shape; more are checked when the truncation
  // keeps more)
  std::map<std::pair<std::array<int64_t, 3>, Charge>, int64_t> svd_kept_hint;
  // A second stream and rocBLAS handle on the same device (created on first
  // use, default priority) for work one call runs on two lanes at once: the
  // leading-values SVD of two large blocks.
  std::unique_ptr<Ctx> aux;
  explicit Ctx(int dev);
  ~Ctx();
  void* blas_handle();
  Ctx& aux_lane() {
    if (!aux) aux.reset(new Ctx(device));
    return *aux;
  }
};

bool indices_match(const QIndex& a, bool conjA, const QIndex& b, bool conjB);

TensorPtr contract(Ctx& ctx, const BSTensor& A, bool conjA, const BSTensor& B, bool conjB,
                   const std::vector<std::pair<int, int>>& pairs);
TensorPtr reorder(Ctx& ctx, const BSTensor& A, bool conjA, const std::vector<int>& order,
                  int numrows);
// Permute every block of A into `out` (already laid out with blocks in the
// permuted sector order) -- used by reorder and contract.
void permute_tensor(Ctx& ctx, const BSTensor& A, const std::vector<int>& order, bool conj,
                    const BSTensor& out_layout, cplx* out);

}  // namespace cam
