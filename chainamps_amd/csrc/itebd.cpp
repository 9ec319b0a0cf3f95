// Device-resident iTEBD.
//
// TrotterDecomposition (TEBD_routines.cpp:14-85): order 1 uses
// {G_odd(tau), G_even(tau)}, order 2 {G_odd(tau/2), G_even(tau), G_odd(tau)},
// order 4 the 11 pieces of Suzuki's product.  iTEBD::evolve
// (TEBD_routines.cpp:120-224) without measurement steps: order 1 applies
// ops 0, 1 per step; order 2 applies op 0, then (1, 2) for all but the last
// step, then 1, 0 and swaps the cell back; order 4 applies piece 0 once,
// then per step pieces 1..9 with OrthogonaliseInversionSymmetric after each
// even-bond piece (the odd positions), and the odd full step 2 that joins the
// step's last half step to the next step's first.
// Each gate is iTEBD::apply_and_decompose (TEBD_routines.cpp:233-266),
// which leaves the cell swapped: A = {A1', U}, lambda = {new, old lambda0}.
#include <chrono>
#include <cmath>

#include "itebd.hpp"
#include "ftebd.hpp"
#include "ortho.hpp"

namespace cam {

ITEBD::ITEBD(Ctx& c, const BSTensor& W, TensorPtr A0, TensorPtr A1, std::vector<double> l0,
             std::vector<double> l1, double time_step, int trotter_order)
    : ctx(&c), order(trotter_order), dt(time_step) {
  CAM_REQUIRE(order == 1 || order == 2 || order == 4, "Only 1st, 2nd and 4th order decompositions are supported");
  CAM_REQUIRE(A0 && A1 && A0->nidx() == 3 && A1->nidx() == 3, "iTEBD: the unit cell needs two 3-index tensors");
  CAM_REQUIRE(!l0.empty() && !l1.empty(), "iTEBD: empty bond weights");
  std::vector<TensorPtr> bond_ops;
  std::vector<int> seq;
  trotter_operators(c, W, dt, order, bond_ops, seq);
  for (int k : seq) ops.push_back(bond_ops.at(k));   // OrderedOperatorPtrs
  initial[0] = A0;
  initial[1] = A1;
  A[0] = std::move(A0);
  A[1] = std::move(A1);
  lam[0] = std::move(l0);
  lam[1] = std::move(l1);
}

double ITEBD::gate(int op, int64_t D, double minS) {
  TensorPtr n0, n1;
  std::vector<double> nl;
  double tr = 0.0;
  tebd_gate(*ctx, *ops.at(op), *A[0], *A[1], lam[0], D, minS, n0, n1, nl, tr);
  A[0] = n0;
  A[1] = n1;
  lam[1] = std::move(lam[0]);
  lam[0] = std::move(nl);
  truncation = tr;
  return tr;
}

void ITEBD::canonise() {
  TensorPtr n0, n1;
  std::vector<double> nl;
  OrthoInfo info;
  orthogonalise_inversion_symmetric(*ctx, A[0], A[1], lam[0], n0, n1, nl, info);
  A[0] = n0;
  A[1] = n1;
  lam[0] = std::move(nl);
}

std::vector<ITEBDMeasure> ITEBD::evolve_measured(int64_t nsteps, int64_t D, double minS, int64_t interval,
                                                 const std::vector<TensorPtr>& measured) {
  CAM_REQUIRE(nsteps >= 0 && interval >= 1, "iTEBD: negative step count or measurement interval below 1");
  std::vector<ITEBDMeasure> out;
  auto swap_cell = [&]() {
    std::swap(A[0], A[1]);
    std::swap(lam[0], lam[1]);
  };
  auto measure = [&](int64_t step) {
    ITEBDMeasure m;
    m.step = step;
    m.time = dt * static_cast<double>(step);
    m.truncation = truncation;
    double s = 0.0;
    for (double x : lam[0])
      if (x > 0.0) s -= x * x * std::log(x * x);
    m.entropy = s;
    m.overlap = overlap(*ctx, {initial[0], initial[1]}, {A[0], A[1]});
    m.abs_overlap = std::hypot(m.overlap.re, m.overlap.im);
    for (const TensorPtr& op : measured) m.values.push_back(one_vertex_measurement(*ctx, *op, A[0], A[1], lam[0]));
    out.push_back(std::move(m));
  };
  auto due = [&](int64_t step) { return step % interval == 0; };
  const int64_t s0 = steps;
  if (order == 1) {
    for (int64_t n = 1; n <= nsteps; ++n) {
      gate(0, D, minS);
      gate(1, D, minS);
      if (due(s0 + n)) {
        canonise();
        measure(s0 + n);
      }
    }
  } else if (order == 2) {
    if (nsteps > 0) {
      gate(0, D, minS);
      for (int64_t n = 1; n < nsteps; ++n) {
        gate(1, D, minS);
        if (due(s0 + n)) {
          gate(0, D, minS);
          swap_cell();
          canonise();
          measure(s0 + n);
          gate(0, D, minS);
        } else {
          gate(2, D, minS);
        }
      }
      gate(1, D, minS);
      gate(0, D, minS);
      swap_cell();
      if (due(s0 + nsteps)) {
        canonise();
        measure(s0 + nsteps);
      }
    }
  } else {
    gate(0, D, minS);
    for (int64_t n = 1; n <= nsteps; ++n) {
      for (int i = 1; i < 10; ++i) {
        gate(i, D, minS);
        if (i % 2) canonise();
      }
      if (due(s0 + n)) {
        gate(10, D, minS);
        swap_cell();
        canonise();
        measure(s0 + n);
        gate(0, D, minS);
      } else {
        gate(2, D, minS);
      }
    }
  }
  steps += nsteps;
  return out;
}

ITEBDRecord ITEBD::evolve(int64_t nsteps, int64_t D, double minS) {
  CAM_REQUIRE(nsteps >= 0, "iTEBD: negative step count");
  const auto t0 = std::chrono::steady_clock::now();
  const double f0 = ctx->contract_flops;
  if (order == 1) {
    for (int64_t s = 0; s < nsteps; ++s) {
      gate(0, D, minS);
      gate(1, D, minS);
    }
  } else if (order == 4) {
    gate(0, D, minS);
    for (int64_t s = 0; s < nsteps; ++s) {
      for (int i = 1; i < 10; ++i) {
        gate(i, D, minS);
        if (i % 2) canonise();
      }
      gate(2, D, minS);
    }
  } else if (nsteps > 0) {
    gate(0, D, minS);
    for (int64_t s = 0; s + 1 < nsteps; ++s) {
      gate(1, D, minS);
      gate(2, D, minS);
    }
    gate(1, D, minS);
    gate(0, D, minS);
    std::swap(A[0], A[1]);   // m_unit.swap(0,1): back to the 12 order
    std::swap(lam[0], lam[1]);
  }
  steps += nsteps;
  ITEBDRecord r{};
  r.steps = steps;
  r.time = dt * static_cast<double>(steps);
  r.truncation = truncation;
  double s = 0.0;
  for (double x : lam[0])
    if (x != 0.0) s -= x * x * std::log(x * x);
  r.entropy = s;
  r.chi = static_cast<int64_t>(lam[0].size());
  r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  r.gate_flops = ctx->contract_flops - f0;
  return r;
}

}  // namespace cam
