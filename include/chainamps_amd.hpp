/*
 * chainamps_amd.hpp -- C++ host layer over the C ABI, mirroring the
 * reference's own interface for the hot path (namespace ajaj in
 * MPX_matrix.hpp, DMRG_routines.hpp, TEBD_routines.hpp) so that a caller of
 * ChainAMPS keeps its call sites: MPX_matrix / MPXIndex / MPXDecomposition,
 * contract(), reorder(), MPX_matrix::SVD(), TwoVertexComponents::HED(),
 * TwoVertexMPOMPSMultiply(), MakeOdd/EvenBondHamiltonian(),
 * MakeBondEvolutionOperator(), iTEBD::evolve() and the iDMRG growth step.
 *
 * Tensors live on the device (one MI355X); host data crosses only where the
 * reference's signatures take host data (CSC construction, eigenvectors,
 * matvec buffers).  Errors follow the reference: the message is printed to
 * std::cout and the process exits with status 1; define CAM_AJAJ_THROW
 * before including this header to get a std::runtime_error instead.
 *
 * Header only; link with -lchainamps_amd.  C++11.
 */
#ifndef CHAINAMPS_AMD_HPP
#define CHAINAMPS_AMD_HPP

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "chainamps_amd.h"

namespace cam_ajaj {

typedef long long MPXInt;
typedef std::pair<MPXInt, MPXInt> MPXPair;   // (index of A, index of B) contracted together
typedef std::vector<short> State;            // one charge per rule (states.hpp:32-61)
typedef std::vector<State> StateArray;

inline void check(int rc) {
  if (rc == 0) return;
  const std::string msg = cam_last_error();
#ifdef CAM_AJAJ_THROW
  throw std::runtime_error(msg);
#else
  std::cout << msg << std::endl;
  std::exit(1);
#endif
}

/* One device context per process thread (the reference is single-threaded on
 * the host side). */
class Context {
 public:
  static cam_ctx* get(int device = 0) {
    static Context c(device);
    return c.ctx_;
  }
  ~Context() {
    if (ctx_) cam_ctx_destroy(ctx_);
  }

 private:
  explicit Context(int device) { check(cam_ctx_create(device, &ctx_)); }
  cam_ctx* ctx_ = nullptr;
};

/* MPXIndex (MPXIndex.hpp:17-65): direction, physical flag, charges. */
class MPXIndex {
 public:
  MPXIndex(bool inward, const StateArray& states, bool physical = false)
      : inward_(inward), physical_(physical), states_(states) {}
  bool Ingoing() const { return inward_; }
  bool Outgoing() const { return !inward_; }
  bool Physical() const { return physical_; }
  MPXInt size() const { return static_cast<MPXInt>(states_.size()); }
  const State& at(MPXInt i) const { return states_.at(static_cast<size_t>(i)); }
  const StateArray& states() const { return states_; }
  MPXIndex flip() const { return MPXIndex(!inward_, states_, physical_); }

 private:
  bool inward_, physical_;
  StateArray states_;
};

class MPX_matrix;
struct MPXDecomposition;
MPX_matrix contract(const MPX_matrix& A, bool conjA, const MPX_matrix& B, bool conjB,
                    const std::vector<MPXPair>& contractidxs);

/* MPX_matrix (MPX_matrix.hpp:61-147): a QN-blocked tensor, resident on the
 * device.  `rules` are the charge rules (states.hpp:32-34: N > 1 is Z_N,
 * N == 1 trivial, N <= 0 U(1)). */
class MPX_matrix {
 public:
  MPX_matrix() = default;
  /* From the reference's storage: indices left to right, numrowindices of
   * them form the CSC rows (mixed radix, first index fastest); x holds
   * complex values, p has cols+1 entries. */
  MPX_matrix(const std::vector<short>& rules, const std::vector<MPXIndex>& indices, MPXInt numrowindices,
             MPXInt rows, MPXInt cols, const std::vector<long long>& p, const std::vector<long long>& i,
             const std::vector<std::complex<double>>& x) {
    std::vector<cam_index_desc> d;
    std::vector<std::vector<int16_t>> q(indices.size());
    for (size_t k = 0; k < indices.size(); ++k) {
      for (const State& s : indices[k].states()) q[k].insert(q[k].end(), s.begin(), s.end());
      cam_index_desc e;
      e.inward = indices[k].Ingoing();
      e.physical = indices[k].Physical();
      e.size = indices[k].size();
      e.charges = q[k].data();
      d.push_back(e);
    }
    std::vector<int64_t> pp(p.begin(), p.end()), ii(i.begin(), i.end());
    if (ii.empty()) ii.push_back(0);
    std::vector<std::complex<double>> xx(x);
    if (xx.empty()) xx.push_back(0.0);
    std::vector<int16_t> r(rules.begin(), rules.end());
    cam_tensor* t = nullptr;
    check(cam_tensor_from_csc(Context::get(), static_cast<int>(r.size()), r.data(), static_cast<int>(d.size()),
                              d.data(), static_cast<int>(numrowindices), rows, cols, pp.data(), ii.data(),
                              reinterpret_cast<const double*>(xx.data()), &t));
    reset(t);
  }
  /* Dense column-major (rows x cols) values, exact zeros not stored. */
  static MPX_matrix FromDense(const std::vector<short>& rules, const std::vector<MPXIndex>& indices,
                              MPXInt numrowindices, const std::vector<std::complex<double>>& dense) {
    MPXInt rows = 1, cols = 1;
    for (MPXInt k = 0; k < numrowindices; ++k) rows *= indices[static_cast<size_t>(k)].size();
    for (size_t k = static_cast<size_t>(numrowindices); k < indices.size(); ++k) cols *= indices[k].size();
    std::vector<long long> p(static_cast<size_t>(cols) + 1, 0), i;
    std::vector<std::complex<double>> x;
    for (MPXInt c = 0; c < cols; ++c) {
      for (MPXInt r = 0; r < rows; ++r) {
        const std::complex<double> v = dense.at(static_cast<size_t>(r + c * rows));
        if (v != 0.0) {
          i.push_back(r);
          x.push_back(v);
        }
      }
      p[static_cast<size_t>(c) + 1] = static_cast<long long>(i.size());
    }
    return MPX_matrix(rules, indices, numrowindices, rows, cols, p, i, x);
  }

  bool isEmpty() const { return !t_; }
  cam_tensor* handle() const { return t_.get(); }
  MPXInt NumRowIndices() const {
    int n = 0, r = 0, nr = 0;
    check(cam_tensor_shape(t_.get(), &n, &r, &nr));
    return r;
  }
  MPXIndex Index(MPXInt k) const {
    int n = 0, r = 0, nr = 0;
    check(cam_tensor_shape(t_.get(), &n, &r, &nr));
    int32_t inw = 0, ph = 0;
    int64_t size = 0;
    check(cam_tensor_index(t_.get(), static_cast<int>(k), &inw, &ph, &size, nullptr));
    std::vector<int16_t> q(static_cast<size_t>(size * nr));
    check(cam_tensor_index(t_.get(), static_cast<int>(k), nullptr, nullptr, nullptr, q.data()));
    StateArray s(static_cast<size_t>(size), State(static_cast<size_t>(nr)));
    for (int64_t e = 0; e < size; ++e)
      for (int c = 0; c < nr; ++c) s[static_cast<size_t>(e)][static_cast<size_t>(c)] = q[static_cast<size_t>(e * nr + c)];
    return MPXIndex(inw != 0, s, ph != 0);
  }
  std::vector<MPXInt> dimsvector() const {
    int n = 0, r = 0, nr = 0;
    check(cam_tensor_shape(t_.get(), &n, &r, &nr));
    std::vector<MPXInt> d;
    for (int k = 0; k < n; ++k) {
      int64_t size = 0;
      check(cam_tensor_index(t_.get(), k, nullptr, nullptr, &size, nullptr));
      d.push_back(size);
    }
    return d;
  }
  /* Host CSC copy (exact zeros dropped). */
  void to_csc(std::vector<long long>& p, std::vector<long long>& i, std::vector<std::complex<double>>& x) const {
    int64_t nnz = 0;
    MPXInt cols = 1;
    const std::vector<MPXInt> d = dimsvector();
    for (size_t k = static_cast<size_t>(NumRowIndices()); k < d.size(); ++k) cols *= d[k];
    std::vector<int64_t> pp(static_cast<size_t>(cols) + 1);
    check(cam_tensor_to_csc(t_.get(), 0, nullptr, nullptr, nullptr, &nnz));   // size only
    std::vector<int64_t> ii(static_cast<size_t>(std::max<int64_t>(nnz, 1)));
    x.assign(static_cast<size_t>(std::max<int64_t>(nnz, 1)), 0.0);
    check(cam_tensor_to_csc(t_.get(), nnz, pp.data(), ii.data(), reinterpret_cast<double*>(x.data()), &nnz));
    p.assign(pp.begin(), pp.end());
    i.assign(ii.begin(), ii.begin() + nnz);
    x.resize(static_cast<size_t>(nnz));
  }

  MPX_matrix& ShiftNumRowIndices(MPXInt numrows) {
    check(cam_tensor_shift_num_row_indices(t_.get(), static_cast<int>(numrows)));
    return *this;
  }
  MPX_matrix& RemoveDummyIndices(const std::vector<MPXInt>& which) {
    std::vector<int64_t> w(which.begin(), which.end());
    check(cam_tensor_remove_dummy_indices(t_.get(), static_cast<int>(w.size()), w.data()));
    return *this;
  }
  MPX_matrix& MoveDummyIndices(const std::vector<MPXInt>& order) {
    std::vector<int64_t> o(order.begin(), order.end());
    check(cam_tensor_move_dummy_indices(t_.get(), static_cast<int>(o.size()), o.data()));
    return *this;
  }
  MPX_matrix& Rescale(std::complex<double> factor) {
    check(cam_tensor_rescale(t_.get(), factor.real(), factor.imag()));
    return *this;
  }
  /* MPX_matrix::store (MPX_matrix.cpp:238-247): the reference's binary
   * .MPX_matrix / .MPS_matrix file.  Returns false on success, as the
   * reference's does. */
  bool store(const std::string& filename) const {
    check(cam_tensor_store(t_.get(), filename.c_str()));
    return false;
  }
  /* MPX_matrix::SVD (MPX_matrix.hpp:116). */
  MPXDecomposition SVD(size_t bond_dimension = 0, double min_s_val = 0.0) const;

  explicit MPX_matrix(cam_tensor* t) { reset(t); }

 private:
  void reset(cam_tensor* t) {
    t_ = std::shared_ptr<cam_tensor>(t, [](cam_tensor* p) {
      if (p) cam_tensor_destroy(p);
    });
  }
  std::shared_ptr<cam_tensor> t_;
};

/* MPXDecomposition (MPX_matrix.hpp:152-184): ColumnMatrix = U,
 * RowMatrix = V^dagger. */
struct MPXDecomposition {
  double Truncation = 0.0;
  MPX_matrix ColumnMatrix;
  std::vector<double> Values;
  MPX_matrix RowMatrix;
  const std::vector<double>& SquareRescale(double sqsum) {
    double w = 0.0;
    for (double v : Values) w += v * v;
    const double f = std::sqrt(sqsum / w);
    for (double& v : Values) v *= f;
    Truncation /= sqsum;
    return Values;
  }
};

inline MPXDecomposition MPX_matrix::SVD(size_t bond_dimension, double min_s_val) const {
  cam_tensor *U = nullptr, *Vd = nullptr;
  int64_t n = 0;
  double disc = 0.0;
  // at most min(rows, cols) values
  const std::vector<MPXInt> dims = dimsvector();
  const MPXInt nr = NumRowIndices();
  MPXInt rows = 1, cols = 1;
  for (size_t k = 0; k < dims.size(); ++k) (static_cast<MPXInt>(k) < nr ? rows : cols) *= dims[k];
  std::vector<double> vals(static_cast<size_t>(std::max<MPXInt>(std::min(rows, cols), 1)));
  check(cam_svd_truncate(Context::get(), t_.get(), static_cast<int64_t>(bond_dimension), min_s_val, &U, &Vd,
                         vals.data(), static_cast<int64_t>(vals.size()), &n, &disc));
  MPXDecomposition d;
  d.Truncation = disc;
  d.ColumnMatrix = MPX_matrix(U);
  d.RowMatrix = MPX_matrix(Vd);
  d.Values.assign(vals.begin(), vals.begin() + n);
  return d;
}

/* ajaj::contract (MPX_matrix.hpp:25, MPX_matrix.cpp:744-791). */
inline MPX_matrix contract(const MPX_matrix& A, bool conjA, const MPX_matrix& B, bool conjB,
                           const std::vector<MPXPair>& contractidxs) {
  std::vector<int64_t> flat;
  for (const MPXPair& p : contractidxs) {
    flat.push_back(p.first);
    flat.push_back(p.second);
  }
  cam_tensor* C = nullptr;
  check(cam_contract(Context::get(), A.handle(), conjA, B.handle(), conjB, static_cast<int>(contractidxs.size()),
                     flat.data(), &C));
  return MPX_matrix(C);
}
inline MPX_matrix contract(const MPX_matrix& A, bool conjA, const MPX_matrix& B, bool conjB,
                           const MPXPair& contractidx) {
  return contract(A, conjA, B, conjB, std::vector<MPXPair>(1, contractidx));
}

/* load_MPX_matrix (MPX_matrix.cpp:594-612): physical indices take the
 * charges of the basis states `spectrum` (one State per basis state). */
inline MPX_matrix load_MPX_matrix(const std::string& filename, const std::vector<short>& rules,
                                  const StateArray& spectrum) {
  std::vector<int16_t> r(rules.begin(), rules.end()), q;
  for (const State& s : spectrum) q.insert(q.end(), s.begin(), s.end());
  cam_tensor* t = nullptr;
  check(cam_tensor_load(Context::get(), filename.c_str(), static_cast<int>(r.size()), r.data(),
                        static_cast<int64_t>(spectrum.size()), q.data(), &t));
  return MPX_matrix(t);
}

/* ajaj::reorder (MPX_matrix.cpp:630-638). */
inline MPX_matrix reorder(const MPX_matrix& A, bool conjA, const std::vector<MPXInt>& neworder,
                          MPXInt numrowindices) {
  std::vector<int64_t> o(neworder.begin(), neworder.end());
  cam_tensor* C = nullptr;
  check(cam_reorder(Context::get(), A.handle(), conjA, static_cast<int>(o.size()), o.data(),
                    static_cast<int>(numrowindices), &C));
  return MPX_matrix(C);
}

/* SparseHED as returned by TwoVertexComponents::HED: eigenvalues and the
 * eigenvector in the compressed (allowed_indices) order. */
struct SparseHED {
  std::vector<double> Values;
  std::vector<std::complex<double>> EigenVector;
  double info[4] = {0, 0, 0, 0};   // matvecs fp64, matvecs fp32, restarts, residual
};

/* TwoVertexComponents (DMRG_routines.hpp:189-214). */
class TwoVertexComponents {
 public:
  TwoVertexComponents(const MPX_matrix& LeftBlock, const MPX_matrix& H, const MPX_matrix& RightBlock) {
    cam_heff* h = nullptr;
    int64_t n = 0;
    check(cam_heff_create(Context::get(), LeftBlock.handle(), H.handle(), RightBlock.handle(), &h, &n));
    h_ = std::shared_ptr<cam_heff>(h, [](cam_heff* p) {
      if (p) cam_heff_destroy(p);
    });
    n_ = n;
  }
  MPXInt size() const { return n_; }
  std::vector<long long> allowed_indices() const {
    std::vector<int64_t> a(static_cast<size_t>(n_));
    check(cam_heff_allowed_map(h_.get(), a.data()));
    return std::vector<long long>(a.begin(), a.end());
  }
  /* TwoVertexComponents::HED (DMRG_routines.cpp:835-896): numevals = 1,
   * which = "SR"; initial: guess in compressed order or nullptr;
   * converge = false is the growth/sweep setting (tol 0.1).  mixed selects
   * the FP32-split bulk with FP64 refinement (an addition; false is the
   * reference path). */
  SparseHED HED(MPXInt numevals, const char which[3], const std::vector<std::complex<double>>* initial,
                bool converge, bool mixed = false) const {
    if (numevals != 1 || std::string(which).substr(0, 2) != "SR") {
      const std::string m = "TwoVertexComponents::HED: only the lowest (numevals=1, which=SR) eigenpair";
#ifdef CAM_AJAJ_THROW
      throw std::runtime_error(m);
#else
      std::cout << m << std::endl;
      std::exit(1);
#endif
    }
    SparseHED ans;
    ans.EigenVector.assign(static_cast<size_t>(n_), 0.0);
    double e = 0.0;
    check(cam_heff_lowest(h_.get(), initial ? reinterpret_cast<const double*>(initial->data()) : nullptr, converge,
                          0.0, 20, mixed, &e, reinterpret_cast<double*>(ans.EigenVector.data()), ans.info));
    ans.Values.push_back(e);
    return ans;
  }
  cam_heff* handle() const { return h_.get(); }

 private:
  std::shared_ptr<cam_heff> h_;
  MPXInt n_ = 0;
};

/* TwoVertexMPOMPSMultiply (DMRG_routines.cpp:898-974): out = H_eff in, host
 * buffers of length size() in the compressed order. */
inline void TwoVertexMPOMPSMultiply(const TwoVertexComponents* comp, std::complex<double>* in,
                                    std::complex<double>* out) {
  check(cam_heff_apply(comp->handle(), reinterpret_cast<const double*>(in), reinterpret_cast<double*>(out),
                       CAM_MEM_HOST, CAM_FP64));
}

/* MakeOddBondHamiltonian / MakeEvenBondHamiltonian / MakeBondEvolutionOperator
 * (TEBD_routines.cpp:826-880). */
inline MPX_matrix MakeOddBondHamiltonian(const MPX_matrix& H) {
  cam_tensor* t = nullptr;
  check(cam_bond_hamiltonian(Context::get(), H.handle(), 1, &t));
  return MPX_matrix(t);
}
inline MPX_matrix MakeEvenBondHamiltonian(const MPX_matrix& H) {
  cam_tensor* t = nullptr;
  check(cam_bond_hamiltonian(Context::get(), H.handle(), 0, &t));
  return MPX_matrix(t);
}
inline MPX_matrix MakeBondEvolutionOperator(const MPX_matrix& BondH, double timestep) {
  cam_tensor* t = nullptr;
  check(cam_bond_evolution_operator(Context::get(), BondH.handle(), timestep, &t));
  return MPX_matrix(t);
}

/* iTEBD (TEBD_routines.hpp): unit cell {A0, A1, Lambdas[0], Lambdas[1]}. */
class iTEBD {
 public:
  iTEBD(const MPX_matrix& H, const MPX_matrix& A0, const MPX_matrix& A1, const std::vector<double>& lambda0,
        const std::vector<double>& lambda1, double time_step_size, int order = 2) {
    cam_itebd* it = nullptr;
    check(cam_itebd_create(Context::get(), H.handle(), A0.handle(), A1.handle(), lambda0.data(),
                           static_cast<int64_t>(lambda0.size()), lambda1.data(),
                           static_cast<int64_t>(lambda1.size()), time_step_size, order, &it));
    it_ = std::shared_ptr<cam_itebd>(it, [](cam_itebd* p) {
      if (p) cam_itebd_destroy(p);
    });
  }
  /* iTEBD::evolve without measurement steps; returns {steps, time,
   * truncation, entropy, chi, seconds, gate flops}. */
  std::vector<double> evolve(MPXInt num_steps, MPXInt bond_dimension, double minS) {
    std::vector<double> r(7);
    check(cam_itebd_evolve(it_.get(), num_steps, bond_dimension, minS, r.data()));
    return r;
  }
  /* iTEBD::evolve(num_steps, measuredMPOs, bond_dimension, minS,
   * measurement_interval) (TEBD_routines.cpp:120-231): one row per
   * measurement, {step, time, truncation, entropy, |overlap|, overlap re,
   * overlap im, then (re, im) of OneVertexMeasurement of each operator}. */
  std::vector<std::vector<double>> evolve(MPXInt num_steps, const std::vector<MPX_matrix>& measuredMPOs,
                                          MPXInt bond_dimension, double minS, MPXInt measurement_interval) {
    std::vector<const cam_tensor*> ops;
    for (const MPX_matrix& m : measuredMPOs) ops.push_back(m.handle());
    const int64_t per = 7 + 2 * static_cast<int64_t>(ops.size());
    const int64_t cap = num_steps / std::max<MPXInt>(measurement_interval, 1) + 1;
    std::vector<double> flat(static_cast<size_t>(cap * per));
    int64_t n = 0;
    check(cam_itebd_evolve_measured(it_.get(), num_steps, bond_dimension, minS, measurement_interval,
                                    static_cast<int>(ops.size()), ops.data(), flat.data(), cap, &n));
    std::vector<std::vector<double>> rows;
    for (int64_t k = 0; k < std::min(n, cap); ++k)
      rows.emplace_back(flat.begin() + k * per, flat.begin() + (k + 1) * per);
    return rows;
  }
  std::vector<double> Lambda(int k) const {
    int64_t n = 0;
    check(cam_itebd_lambda(it_.get(), k, nullptr, 0, &n));
    std::vector<double> v(static_cast<size_t>(std::max<int64_t>(n, 1)));
    check(cam_itebd_lambda(it_.get(), k, v.data(), static_cast<int64_t>(v.size()), &n));
    v.resize(static_cast<size_t>(n));
    return v;
  }
  MPX_matrix Matrix(int k) const {
    cam_tensor* t = nullptr;
    check(cam_itebd_tensor(it_.get(), k, &t));
    return MPX_matrix(t);
  }

 private:
  std::shared_ptr<cam_itebd> it_;
};

/* Finite-chain TEBD (TEBD::evolve, TEBD_routines.cpp:672-791) on an MPS held
 * in device memory: `sites` is a left-canonical MPS of an even number of
 * vertices, each MPS_matrix [s, a_l, a_r]; Trotter order 1, 2 or 4. */
struct TEBDMeasurement {
  MPXInt TimeStep = 0;
  double Time = 0.0, Truncation = 0.0, Entropy = 0.0;
  std::complex<double> Overlap;
};
class TEBD {
 public:
  TEBD(const MPX_matrix& H, const std::vector<MPX_matrix>& sites, double time_step_size, int order = 2) {
    std::vector<const cam_tensor*> h;
    for (const MPX_matrix& m : sites) h.push_back(m.handle());
    cam_ftebd* t = nullptr;
    check(cam_ftebd_create(Context::get(), H.handle(), h.data(), static_cast<int64_t>(h.size()), time_step_size,
                           order, &t));
    t_ = std::shared_ptr<cam_ftebd>(t, [](cam_ftebd* p) {
      if (p) cam_ftebd_destroy(p);
    });
    n_ = static_cast<MPXInt>(sites.size());
  }
  void evolve(MPXInt num_steps, MPXInt bond_dimension, double minS, MPXInt measurement_interval = 1) {
    check(cam_ftebd_evolve(t_.get(), num_steps, bond_dimension, minS, measurement_interval));
  }
  /* one entry per measurement (left_canonise_measure) so far */
  std::vector<TEBDMeasurement> results() const {
    int64_t n = 0;
    check(cam_ftebd_records(t_.get(), nullptr, 0, &n));
    std::vector<double> r(static_cast<size_t>(7 * std::max<int64_t>(n, 1)));
    check(cam_ftebd_records(t_.get(), r.data(), n, &n));
    std::vector<TEBDMeasurement> out(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
      const double* x = r.data() + 7 * i;
      out[i].TimeStep = static_cast<MPXInt>(x[0]);
      out[i].Time = x[1];
      out[i].Truncation = x[2];
      out[i].Entropy = x[3];
      out[i].Overlap = std::complex<double>(x[4], x[5]);
    }
    return out;
  }
  /* left-canonical site v = 1..N after the last step */
  MPX_matrix Matrix(MPXInt v) const {
    cam_tensor* t = nullptr;
    check(cam_ftebd_site(t_.get(), v, &t));
    return MPX_matrix(t);
  }
  MPXInt size() const { return n_; }

 private:
  std::shared_ptr<cam_ftebd> t_;
  MPXInt n_ = 0;
};

class UnitCell;

/* iDMRG growth (iDMRG::run / SuperBlock::grow_two_vertex) and finite
 * steps (SuperBlock::move_right/left_two_vertex): one Energies.dat row
 * (Energy, Energy/vertices, Entropy, Truncation, Fidelity) and the centre's
 * position. */
struct iDMRGStep {
  MPXInt N = 0, chi = 0, n = 0, left = 0, right = 0;
  double Energy = 0.0, EnergyPerVertex = 0.0, Entropy = 0.0, Truncation = 0.0, Fidelity = 0.0,
         Seconds = 0.0;
  static iDMRGStep from_record(const double* r) {
    iDMRGStep s;
    s.N = static_cast<MPXInt>(r[0]);
    s.Energy = r[1];
    s.EnergyPerVertex = r[2];
    s.Entropy = r[3];
    s.Truncation = r[4];
    s.chi = static_cast<MPXInt>(r[5]);
    s.n = static_cast<MPXInt>(r[6]);
    s.Seconds = r[10];
    s.Fidelity = r[15];
    s.left = static_cast<MPXInt>(r[17]);
    s.right = static_cast<MPXInt>(r[18]);
    return s;
  }
};
class iDMRG {
 public:
  iDMRG(const MPX_matrix& H, const State& target, bool converge = false, bool mixed = false) {
    std::vector<int16_t> t(target.begin(), target.end());
    cam_idmrg* d = nullptr;
    check(cam_idmrg_create(Context::get(), H.handle(), t.data(), converge, mixed, 20, &d));
    d_ = std::shared_ptr<cam_idmrg>(d, [](cam_idmrg* p) {
      if (p) cam_idmrg_destroy(p);
    });
  }
  /* UnitCell(getCentralDecomposition(), getPreviousLambda()) of the last
   * growth step (iDMRG_DRV.cpp:70). */
  inline UnitCell CentralUnitCell() const;
  iDMRGStep grow(MPXInt chi, double truncation = 0.0) {
    double r[CAM_IDMRG_RECORD];
    check(cam_idmrg_grow(d_.get(), chi, truncation, r));
    return iDMRGStep::from_record(r);
  }
  /* SuperBlock::move_right_two_vertex / move_left_two_vertex
   * (DMRG_routines.cpp:376-444) on the grown chain */
  iDMRGStep move_right_two_vertex(MPXInt chi, double truncation = 0.0) { return move(true, chi, truncation); }
  iDMRGStep move_left_two_vertex(MPXInt chi, double truncation = 0.0) { return move(false, chi, truncation); }

 protected:
  std::shared_ptr<cam_idmrg> d_;

 private:
  iDMRGStep move(bool right, MPXInt chi, double truncation) {
    double r[CAM_IDMRG_RECORD];
    check(cam_idmrg_move(d_.get(), right, chi, truncation, r));
    return iDMRGStep::from_record(r);
  }
};

/* FiniteDMRG (fDMRG_routines.hpp): growth to the chain length with
 * iDMRG, then FiniteDMRG::run (fDMRG_routines.cpp:19-66), which sweeps the
 * centre right to the end, left to the other end and back to the middle;
 * run() returns the rows the reference writes to Energies.dat (the centre
 * in the middle, on the way left and on the way back). */
class FiniteDMRG : public iDMRG {
 public:
  FiniteDMRG(const MPX_matrix& H, const State& target, bool converge = false, bool mixed = false)
      : iDMRG(H, target, converge, mixed) {}
  std::vector<iDMRGStep> run(MPXInt num_sweeps, MPXInt chi, double truncation = 0.0) {
    int64_t n = 0;
    check(cam_fdmrg_sweeps(d_.get(), num_sweeps, chi, truncation, nullptr, nullptr, 0, &n));
    std::vector<double> r(static_cast<size_t>(n > 0 ? n : 1) * CAM_IDMRG_RECORD);
    std::vector<int32_t> mid(static_cast<size_t>(n > 0 ? n : 1));
    check(cam_fdmrg_sweeps(d_.get(), num_sweeps, chi, truncation, r.data(), mid.data(), n, &n));
    std::vector<iDMRGStep> out;
    for (int64_t i = 0; i < n; ++i)
      if (mid[i]) out.push_back(iDMRGStep::from_record(r.data() + i * CAM_IDMRG_RECORD));
    return out;
  }
};

/* UnitCell (UnitCell.hpp:29-66) on the device: Matrices [s, a_l, a_r] and
 * the lambda on the left of each. */
class UnitCell {
 public:
  std::vector<MPX_matrix> Matrices;
  std::vector<std::vector<double>> Lambdas;
  UnitCell() = default;
  UnitCell(std::vector<MPX_matrix> m, std::vector<std::vector<double>> l) : Matrices(std::move(m)), Lambdas(std::move(l)) {}
  size_t size() const { return Matrices.size(); }
  /* UnitCell::Entropy (UnitCell.cpp:98-106): entropy of Lambdas[0]. */
  double Entropy() const {
    if (Lambdas.empty()) return -1.0;
    double s = 0.0;
    for (double v : Lambdas[0])
      if (v > 0.0) s -= v * v * std::log(v * v);
    return s;
  }
  /* handle over the same device tensors (cam_unitcell_create) */
  std::shared_ptr<cam_unitcell> handle() const {
    std::vector<const cam_tensor*> h;
    std::vector<const double*> l;
    std::vector<int64_t> n;
    for (const MPX_matrix& m : Matrices) h.push_back(m.handle());
    for (const auto& v : Lambdas) {
      l.push_back(v.data());
      n.push_back(static_cast<int64_t>(v.size()));
    }
    cam_unitcell* u = nullptr;
    check(cam_unitcell_create(Context::get(), static_cast<int>(h.size()), h.data(), l.data(), n.data(), &u));
    return std::shared_ptr<cam_unitcell>(u, [](cam_unitcell* p) {
      if (p) cam_unitcell_destroy(p);
    });
  }
  static UnitCell FromHandle(cam_unitcell* u) {
    std::shared_ptr<cam_unitcell> keep(u, [](cam_unitcell* p) {
      if (p) cam_unitcell_destroy(p);
    });
    int n = 0;
    check(cam_unitcell_size(u, &n));
    UnitCell c;
    for (int k = 0; k < n; ++k) {
      cam_tensor* t = nullptr;
      check(cam_unitcell_matrix(u, k, &t));
      c.Matrices.push_back(MPX_matrix(t));
      int64_t len = 0;
      check(cam_unitcell_lambda(u, k, nullptr, 0, &len));
      std::vector<double> v(static_cast<size_t>(len));
      check(cam_unitcell_lambda(u, k, v.data(), len, &len));
      c.Lambdas.push_back(std::move(v));
    }
    return c;
  }
};

inline UnitCell iDMRG::CentralUnitCell() const {
  cam_unitcell* u = nullptr;
  check(cam_idmrg_unitcell(d_.get(), &u));
  return UnitCell::FromHandle(u);
}

/* Orthogonalise (measurement.cpp:180-246) of a two-vertex cell. */
inline UnitCell Orthogonalise(const UnitCell& C) {
  cam_unitcell* out = nullptr;
  check(cam_orthogonalise(C.handle().get(), &out, nullptr));
  return UnitCell::FromHandle(out);
}

/* OrthogonaliseInversionSymmetric (measurement.cpp:131-178). */
inline UnitCell OrthogonaliseInversionSymmetric(const UnitCell& C) {
  cam_unitcell* out = nullptr;
  check(cam_orthogonalise_inversion_symmetric(C.handle().get(), &out, nullptr));
  return UnitCell::FromHandle(out);
}

/* iTwoVertexEnergy (measurement.cpp:371-397): energy per vertex of an
 * orthogonalised two-vertex cell. */
inline std::complex<double> iTwoVertexEnergy(const MPX_matrix& H_MPO, const UnitCell& Ortho) {
  double e[2];
  check(cam_itwovertex_energy(H_MPO.handle(), Ortho.handle().get(), e, nullptr));
  return {e[0], e[1]};
}

}  // namespace cam_ajaj

#endif
