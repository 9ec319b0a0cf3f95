// Device-resident iTEBD time evolution (iTEBD::evolve mirror).
#pragma once
#include <vector>

#include "solvers.hpp"

namespace cam {

struct ITEBDRecord {
  int64_t steps;        // time steps taken so far
  double time;          // steps * dt
  double truncation;    // of the last gate
  double entropy;       // of lambda0 (sum over nonzero -l^2 log l^2)
  int64_t chi;          // length of lambda0
  double seconds;       // wall time of this evolve call
  double gate_flops;    // algorithmic flops of the gate contractions of this call
};

// TrotterDecomposition (TEBD_routines.cpp:14-85) for orders 1, 2 and 4 and
// the evolve loop (TEBD_routines.cpp:120-224) without measurement steps; the
// unit cell {A0, A1, lambda0, lambda1} stays on the device.
// do_measurements (TEBD_routines.cpp:99-118) at a measurement step
struct ITEBDMeasure {
  int64_t step;
  double time, truncation, entropy, abs_overlap;
  cplx overlap;                  // Overlap(initial cell, canonised cell)
  std::vector<cplx> values;      // OneVertexMeasurement of each measured operator
};

struct ITEBD {
  Ctx* ctx;
  int order;
  double dt;
  std::vector<TensorPtr> ops;   // OrderedOperatorPtrs
  TensorPtr A[2];
  std::vector<double> lam[2];
  TensorPtr initial[2];          // m_initial_unit, for the overlap
  int64_t steps = 0;
  double truncation = 0.0;

  ITEBD(Ctx& c, const BSTensor& W, TensorPtr A0, TensorPtr A1, std::vector<double> l0, std::vector<double> l1,
        double time_step, int trotter_order);
  // one apply_and_decompose with the unit-cell swap
  double gate(int op, int64_t D, double minS);
  // OrthogonaliseInversionSymmetric of the cell (order 4, between pieces)
  void canonise();
  ITEBDRecord evolve(int64_t nsteps, int64_t D, double minS);
  // evolve with measurement steps every `interval` steps (of the global
  // step count): the reference's measuring branches, canonisation by
  // OrthogonaliseInversionSymmetric and do_measurements
  std::vector<ITEBDMeasure> evolve_measured(int64_t nsteps, int64_t D, double minS, int64_t interval,
                                            const std::vector<TensorPtr>& measured);
};

}  // namespace cam
