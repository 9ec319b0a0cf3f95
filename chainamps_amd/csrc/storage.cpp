// The reference's binary formats, for checkpoint / resume interop.
//
// MPX_matrix (.MPX_matrix, .MPS_matrix; MPX_matrix::fprint_binary and
// load_MPX_matrix_binary, MPX_matrix.cpp:219-247, 594-612):
//   size_t   number of indices
//   int64    number of row indices
//   per index (MPXIndex::fprint_binary, MPXIndex.cpp:8-24; load 98-120):
//     bool inward, bool physical, size_t size,
//     and for a non-physical index `size` states of nrules int16 charges
//     (State::fprint_binary, states.cpp:59-61); a physical index stores no
//     charges (the reader takes them from the basis)
//   the CSC matrix (SparseMatrix::fprint_binary, sparse_interface.cpp:727-764):
//     int64 nzmax, nz (= -1 for CSC), rows, cols; int64 p[cols + 1];
//     then int64 i[nzmax], complex<double> x[nzmax] -- the reader
//     (load_SparseMatrix_binary, 768-835) reads nothing after p when
//     nzmax == 0 or p[cols] == 0; this writer stores nzmax = 0 for an empty
//     matrix, which both readers accept.
// Basis (Basis::fprint_binary / load_Basis_binary, states.cpp:119-189):
//   size_t nrules, int16 rules[nrules], size_t d, int16 charges[d][nrules],
//   double energies[d].
// UnitCell (.UNITCELL; UnitCell::store / load_UnitCell_binary,
// UnitCell.cpp:70-95, 112-145): the basis, size_t n, n MPX_matrix records,
// then n times (size_t length, double lambda[length]).
// Integers are the host's (little-endian, 64-bit size_t), as the reference
// writes them.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "storage.hpp"

namespace cam {

namespace {

struct File {
  std::FILE* f = nullptr;
  std::string path;
  File(const std::string& p, const char* mode) : path(p) {
    f = std::fopen(p.c_str(), mode);
    CAM_REQUIRE(f != nullptr, "cannot open " + p);
  }
  ~File() {
    if (f) std::fclose(f);
  }
  void put(const void* d, size_t n) { CAM_REQUIRE(std::fwrite(d, 1, n, f) == n, "write failed: " + path); }
  void get(void* d, size_t n) { CAM_REQUIRE(std::fread(d, 1, n, f) == n, "unexpected end of file: " + path); }
  template <typename T> void put(const T& v) { put(&v, sizeof(T)); }
  template <typename T> T get() {
    T v;
    get(&v, sizeof(T));
    return v;
  }
};

void write_mpx(File& F, const BSTensor& t, hipStream_t s) {
  const int nr = t.rules.n();
  const size_t nidx = static_cast<size_t>(t.nidx());
  F.put<size_t>(nidx);
  F.put<int64_t>(t.numrows);
  for (const QIndex& q : t.idx) {
    F.put<bool>(q.inward);
    F.put<bool>(q.physical);
    F.put<size_t>(static_cast<size_t>(q.size));
    if (!q.physical)
      for (int64_t e = 0; e < q.size; ++e) F.put(q.charges[e].q.data(), sizeof(int16_t) * nr);
  }
  const int64_t rows = t.rows(), cols = t.cols();
  const int64_t nnz = tensor_to_csc(t, 0, nullptr, nullptr, nullptr, s);
  std::vector<int64_t> p(cols + 1), i(std::max<int64_t>(nnz, 1));
  std::vector<double> x(2 * std::max<int64_t>(nnz, 1));
  tensor_to_csc(t, nnz, p.data(), i.data(), x.data(), s);
  const int64_t params[4] = {nnz, -1, rows, cols};
  F.put(params, sizeof(params));
  F.put(p.data(), sizeof(int64_t) * (cols + 1));
  if (nnz > 0) {
    F.put(i.data(), sizeof(int64_t) * nnz);
    F.put(x.data(), sizeof(double) * 2 * nnz);
  }
}

TensorPtr read_mpx(File& F, const Rules& R, const std::vector<Charge>& phys, hipStream_t s) {
  const size_t nidx = F.get<size_t>();
  const int64_t numrows = F.get<int64_t>();
  CAM_REQUIRE(nidx >= 1 && nidx <= 16, "MPX_matrix file: unsupported number of indices");
  std::vector<QIndex> idx(nidx);
  const int nr = R.n();
  for (QIndex& q : idx) {
    q.inward = F.get<bool>();
    q.physical = F.get<bool>();
    q.size = static_cast<int64_t>(F.get<size_t>());
    CAM_REQUIRE(q.size >= 0, "MPX_matrix file: bad index size");
    if (q.physical) {
      CAM_REQUIRE(static_cast<int64_t>(phys.size()) == q.size,
                  "MPX_matrix file: physical index size differs from the basis");
      q.charges = phys;
    } else {
      q.charges.resize(q.size);
      std::vector<int16_t> buf(nr);
      for (int64_t e = 0; e < q.size; ++e) {
        F.get(buf.data(), sizeof(int16_t) * nr);
        q.charges[e] = R.normalise(buf.data());
      }
    }
  }
  int64_t params[4];
  F.get(params, sizeof(params));
  const int64_t nzmax = params[0], nz = params[1], rows = params[2], cols = params[3];
  CAM_REQUIRE(nz == -1, "MPX_matrix file: matrix not in CSC form");
  CAM_REQUIRE(rows >= 0 && cols >= 0 && nzmax >= 0, "MPX_matrix file: bad matrix header");
  std::vector<int64_t> p(cols + 1);
  F.get(p.data(), sizeof(int64_t) * (cols + 1));
  const int64_t nnz = p[cols];
  std::vector<int64_t> i(std::max<int64_t>(nzmax, 1));
  std::vector<double> x(2 * std::max<int64_t>(nzmax, 1));
  if (nzmax > 0 && nnz > 0) {
    CAM_REQUIRE(nnz <= nzmax, "MPX_matrix file: column pointers exceed nzmax");
    F.get(i.data(), sizeof(int64_t) * nzmax);
    F.get(x.data(), sizeof(double) * 2 * nzmax);
  }
  return tensor_from_csc(R, std::move(idx), static_cast<int>(numrows), rows, cols, p.data(), i.data(), x.data(),
                         s);
}

void write_basis(File& F, const StoredBasis& b) {
  F.put<size_t>(b.rules.N.size());
  F.put(b.rules.N.data(), sizeof(int16_t) * b.rules.N.size());
  F.put<size_t>(b.charges.size());
  const int nr = b.rules.n();
  for (const Charge& c : b.charges) F.put(c.q.data(), sizeof(int16_t) * nr);
  F.put(b.energies.data(), sizeof(double) * b.energies.size());
}

StoredBasis read_basis(File& F) {
  StoredBasis b;
  const size_t nr = F.get<size_t>();
  CAM_REQUIRE(nr >= 1 && nr <= static_cast<size_t>(kMaxRules), "basis file: unsupported number of charge rules");
  b.rules.N.resize(nr);
  F.get(b.rules.N.data(), sizeof(int16_t) * nr);
  const size_t d = F.get<size_t>();
  b.charges.resize(d);
  std::vector<int16_t> buf(nr);
  for (size_t e = 0; e < d; ++e) {
    F.get(buf.data(), sizeof(int16_t) * nr);
    b.charges[e] = b.rules.normalise(buf.data());
  }
  b.energies.resize(d);
  F.get(b.energies.data(), sizeof(double) * d);
  return b;
}

}  // namespace

void store_tensor(const std::string& path, const BSTensor& t, hipStream_t s) {
  File F(path, "wb");
  write_mpx(F, t, s);
}

TensorPtr load_tensor(const std::string& path, const Rules& rules, const std::vector<Charge>& phys, hipStream_t s) {
  File F(path, "rb");
  return read_mpx(F, rules, phys, s);
}

void store_unitcell(const std::string& path, const StoredBasis& basis, const std::vector<TensorPtr>& matrices,
                    const std::vector<std::vector<double>>& lambdas, hipStream_t s) {
  CAM_REQUIRE(matrices.size() == lambdas.size(), "Invalid Unit Cell: matrices and lambdas differ in number");
  CAM_REQUIRE(basis.energies.size() == basis.charges.size(), "basis: charges and energies differ in number");
  File F(path, "wb");
  write_basis(F, basis);
  F.put<size_t>(matrices.size());
  for (const TensorPtr& m : matrices) write_mpx(F, *m, s);
  for (const auto& l : lambdas) {
    F.put<size_t>(l.size());
    F.put(l.data(), sizeof(double) * l.size());
  }
}

void load_unitcell(const std::string& path, StoredBasis& basis, std::vector<TensorPtr>& matrices,
                   std::vector<std::vector<double>>& lambdas, hipStream_t s) {
  File F(path, "rb");
  basis = read_basis(F);
  const size_t n = F.get<size_t>();
  CAM_REQUIRE(n <= 1024, "UNITCELL file: implausible number of matrices");
  matrices.clear();
  lambdas.clear();
  for (size_t k = 0; k < n; ++k) matrices.push_back(read_mpx(F, basis.rules, basis.charges, s));
  for (size_t k = 0; k < n; ++k) {
    const size_t len = F.get<size_t>();
    CAM_REQUIRE(len <= (size_t(1) << 32), "UNITCELL file: implausible lambda length");
    std::vector<double> l(len);
    F.get(l.data(), sizeof(double) * len);
    lambdas.push_back(std::move(l));
  }
}

}  // namespace cam
