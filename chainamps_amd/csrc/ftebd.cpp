// Finite-chain TEBD (TEBD_routines.cpp:273-791) on a device-resident MPS.
//
// The reference keeps every site matrix on disk ("Evolving_<name>_Left_v"
// and "..._Right_k" files, k = N + 1 - v counting from the right end); here
// site v's left-shaped tensor [s, a_l, a_r] (two row indices) and
// right-shaped tensor [a_l, s, a_r] (one) stay in HBM and each bond update
// is the gate contraction and blockwise SVD of the iTEBD path:
//   apply_to_odd_bonds (273-318): for v = N, N-2, ..., 2
//     theta = reorder(contract(G, contract(L[v-1], R, 21), 2032), 0213, 2)
//     theta = U S V^dagger (D, minS);  right[v] = V^dagger
//     right_shape(U S) = C s' W^dagger (untruncated);  right[v-1] = W^dagger
//     (times the phase Tr C at v = 2), R = contract(L[v-2], C s', 20);
//   apply_to_even_bonds (320-391): the same for v = N-1, N-3, ..., 3 after
//     an untruncated rotation of site N, then a last rotation of site 1;
//   left_canonise[_measure] (395-547): from site 1 rightwards, SVD of
//     left_shape(Vd . right[v]) (no truncation as evolve calls it), values
//     square-rescaled to 1, Vd = diag(s) V^dagger carried on, the phase
//     Tr Vd folded into the last site; when measuring, the entropy of bond
//     N/2, the largest truncation of the step and the overlap with the
//     initial state are recorded;
//   evolve (672-791): Trotter orders 1, 2 and 4, as the reference sequences
//     them, with a measurement every `interval` steps.
#include <cmath>

#include "ftebd.hpp"
#include "kernels.hpp"

namespace cam {

namespace {

struct Decomp {
  TensorPtr U, Vd;
  std::vector<double> values;
  double truncation = 0.0;
};

Decomp svd(Ctx& ctx, const BSTensor& T, int64_t D, double minS) {
  Decomp d;
  svd_truncate(ctx, T, D, minS, d.U, d.Vd, d.values, d.truncation);
  return d;
}

TensorPtr right_shape(Ctx& ctx, const BSTensor& A) { return reorder(ctx, A, false, {1, 0, 2}, 1); }
TensorPtr left_shape(Ctx& ctx, const BSTensor& A) { return reorder(ctx, A, false, {1, 0, 2}, 2); }

TensorPtr copy_of(Ctx& ctx, const BSTensor& A) {
  auto t = std::make_shared<BSTensor>();
  t->rules = A.rules;
  t->idx = A.idx;
  t->numrows = A.numrows;
  t->blocks = A.blocks;
  t->finalize_layout();
  if (A.total)
    HIP_CHECK(hipMemcpyAsync(t->data.get(), A.data.get(), A.total * sizeof(cplx), hipMemcpyDeviceToDevice,
                             ctx.stream));
  return t;
}

// MPX_matrix::Trace of a square 2-index tensor: sum of T[p, p]
cplx trace(Ctx& ctx, const BSTensor& T) {
  CAM_REQUIRE(T.nidx() == 2 && T.idx[0].size == T.idx[1].size, "Trace needs a square 2-index tensor");
  const std::vector<cplx> h = download(T, ctx.stream);
  const QIndex &a = T.idx[0], &b = T.idx[1];
  cplx acc{0.0, 0.0};
  for (int64_t p = 0; p < a.size; ++p) {
    const int k = T.find_block({a.pos_sec[p], b.pos_sec[p]});
    if (k < 0) continue;
    const Block& blk = T.blocks[k];
    const cplx v = h[blk.offset + a.pos_off[p] + b.pos_off[p] * blk.dims[0]];
    acc.re += v.re;
    acc.im += v.im;
  }
  return acc;
}

void rescale(Ctx& ctx, BSTensor& T, cplx f) {
  if (T.total) launch_scale(f, T.data.get(), T.total, ctx.stream);
}

}  // namespace

void trotter_operators(Ctx& ctx, const BSTensor& W, double tau, int order, std::vector<TensorPtr>& ops,
                       std::vector<int>& seq) {
  CAM_REQUIRE(order == 1 || order == 2 || order == 4, "Only 1st, 2nd and 4th order decompositions are supported");
  TensorPtr Ho = bond_hamiltonian(ctx, W, true);
  TensorPtr He = bond_hamiltonian(ctx, W, false);
  ops.clear();
  if (order == 1) {
    ops = {bond_evolution_operator(ctx, *Ho, tau), bond_evolution_operator(ctx, *He, tau)};
    seq = {0, 1};
  } else if (order == 2) {
    ops = {bond_evolution_operator(ctx, *Ho, 0.5 * tau), bond_evolution_operator(ctx, *He, tau),
           bond_evolution_operator(ctx, *Ho, tau)};
    seq = {0, 1, 2};
  } else {
    // Suzuki's fourth-order product: p = 1 / (4 - 4^(1/3)), the middle
    // piece 1 - 4p and the joined odd half steps p + (1 - 4p) = 1 - 3p
    const double p = 1.0 / (4.0 - std::cbrt(4.0));
    const double f3 = 1.0 - 4.0 * p, f13 = 1.0 - 3.0 * p;
    ops = {bond_evolution_operator(ctx, *Ho, 0.5 * tau * p), bond_evolution_operator(ctx, *He, tau * p),
           bond_evolution_operator(ctx, *Ho, tau * p), bond_evolution_operator(ctx, *Ho, 0.5 * tau * f13),
           bond_evolution_operator(ctx, *He, tau * f3)};
    seq = {0, 1, 2, 1, 3, 4, 3, 1, 2, 1, 0};
  }
}

FTEBD::FTEBD(Ctx& c, const BSTensor& W, const std::vector<TensorPtr>& sites, double timestep, int trotter_order)
    : ctx(&c), order(trotter_order), dt(timestep), N(static_cast<int64_t>(sites.size())) {
  CAM_REQUIRE(N >= 2 && N % 2 == 0, "finite TEBD needs an even number of vertices >= 2");
  left.assign(N + 1, nullptr);
  right.assign(N + 1, nullptr);
  initial.assign(N + 1, nullptr);
  for (int64_t v = 1; v <= N; ++v) {
    const TensorPtr& s = sites[v - 1];
    CAM_REQUIRE(s && s->nidx() == 3 && s->idx[0].physical && s->idx[1].inward && !s->idx[2].inward,
                "finite TEBD: sites must be left-shaped MPS matrices [s, a_l (in), a_r (out)]");
    left[v] = copy_of(c, *s);
    left[v]->numrows = 2;
    initial[v] = copy_of(c, *s);
    initial[v]->numrows = 2;
  }
  trotter_operators(c, W, dt, order, ops, seq);
}

void FTEBD::apply_to_odd_bonds(const BSTensor& G, int64_t D, double minS) {
  Ctx& c = *ctx;
  TensorPtr R = left[N];
  for (int64_t v = N; v > 1; v -= 2) {
    TensorPtr X = contract(c, *left[v - 1], false, *R, false, {{2, 1}});
    TensorPtr theta = reorder(c, *contract(c, G, false, *X, false, {{2, 0}, {3, 2}}), false, {0, 2, 1, 3}, 2);
    Decomp d = svd(c, *theta, D, minS);
    max_truncation = std::max(max_truncation, d.truncation);
    right[v] = d.Vd;
    scale_along_index(c, *d.U, 2, d.values);
    Decomp rot = svd(c, *right_shape(c, *d.U), 0, 0.0);
    max_truncation = std::max(max_truncation, rot.truncation);
    if (v == 2) rescale(c, *rot.Vd, trace(c, *rot.U));
    right[v - 1] = rot.Vd;
    if (v > 2) {
      scale_along_index(c, *rot.U, 1, rot.values);
      R = contract(c, *left[v - 2], false, *rot.U, false, {{2, 0}});
    }
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

void FTEBD::apply_to_even_bonds(const BSTensor& G, int64_t D, double minS) {
  Ctx& c = *ctx;
  Decomp rot = svd(c, *right_shape(c, *left[N]), 0, 0.0);
  right[N] = rot.Vd;
  scale_along_index(c, *rot.U, 1, rot.values);
  TensorPtr R = contract(c, *left[N - 1], false, *rot.U, false, {{2, 0}});
  for (int64_t v = N - 1; N > 2 && v > 1; v -= 2) {
    TensorPtr X = contract(c, *left[v - 1], false, *R, false, {{2, 1}});
    TensorPtr theta = reorder(c, *contract(c, G, false, *X, false, {{2, 0}, {3, 2}}), false, {0, 2, 1, 3}, 2);
    Decomp d = svd(c, *theta, D, minS);
    max_truncation = std::max(max_truncation, d.truncation);
    right[v] = d.Vd;
    scale_along_index(c, *d.U, 2, d.values);
    Decomp r2 = svd(c, *right_shape(c, *d.U), 0, 0.0);
    max_truncation = std::max(max_truncation, r2.truncation);
    right[v - 1] = r2.Vd;
    if (v > 2) {
      scale_along_index(c, *r2.U, 1, r2.values);
      R = contract(c, *left[v - 2], false, *r2.U, false, {{2, 0}});
    }
  }
  Decomp last = svd(c, *right_shape(c, *R), 0, 0.0);
  rescale(c, *last.Vd, trace(c, *last.U));
  right[1] = last.Vd;
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

void FTEBD::left_canonise(bool measure) {
  Ctx& c = *ctx;
  FTEBDRecord rec{};
  rec.step = step;
  rec.time = static_cast<double>(step) * dt;
  auto square_rescale = [](std::vector<double>& vals) {
    double sq = 0.0;
    for (double x : vals) sq += x * x;
    CAM_REQUIRE(sq > 0.0, "Square sum error");
    for (double& x : vals) x *= std::sqrt(1.0 / sq);
  };
  Decomp first = svd(c, *left_shape(c, *right[1]), 0, 0.0);
  square_rescale(first.values);
  if (measure && N == 2) {
    rec.truncation = max_truncation;
    rec.entropy = entropy(first.values);
  }
  left[1] = first.U;
  TensorPtr overlap;
  if (measure) overlap = contract(c, *initial[1], true, *first.U, false, {{0, 0}, {1, 1}});
  TensorPtr Vd = first.Vd;
  scale_along_index(c, *Vd, 0, first.values);
  for (int64_t v = 2; v <= N; ++v) {
    TensorPtr X = contract(c, *Vd, false, *right[v], false, {{1, 0}});   // [k, s, a_r]
    Decomp d = svd(c, *left_shape(c, *X), 0, 0.0);
    square_rescale(d.values);
    if (measure && v == N / 2) {
      rec.truncation = max_truncation;
      rec.entropy = entropy(d.values);
    }
    Vd = d.Vd;
    scale_along_index(c, *Vd, 0, d.values);
    if (v == N) rescale(c, *d.U, trace(c, *Vd));
    if (measure)
      overlap = contract(c, *initial[v], true, *contract(c, *overlap, false, *d.U, false, {{1, 1}}), false,
                         {{0, 1}, {1, 0}});
    left[v] = d.U;
  }
  if (measure) {
    const cplx ov = trace(c, *overlap);
    rec.overlap_re = ov.re;
    rec.overlap_im = ov.im;
    rec.abs_overlap = std::hypot(ov.re, ov.im);
    records.push_back(rec);
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

void FTEBD::evolve(int64_t steps, int64_t D, double minS, int64_t interval) {
  CAM_REQUIRE(steps >= 0 && interval >= 1, "finite TEBD: steps >= 0 and interval >= 1");
  auto bonds = [&](int k) {
    const BSTensor& G = *ops.at(seq.at(k));
    if (k % 2 == 0)
      apply_to_odd_bonds(G, D, minS);
    else
      apply_to_even_bonds(G, D, minS);
  };
  if (order == 1) {
    for (int64_t n = 0; n < steps; ++n) {
      ++step;
      bonds(0);
      if (N > 2) {
        left_canonise(false);
        bonds(1);
      }
      left_canonise(step % interval == 0);
      max_truncation = 0.0;
    }
  } else if (order == 2) {
    if (steps > 0) {
      bonds(0);
      left_canonise(false);
      for (int64_t n = 0; n < steps - 1; ++n) {
        ++step;
        bonds(1);
        left_canonise(false);
        if (step % interval == 0) {
          bonds(0);
          left_canonise(true);
          bonds(0);
          left_canonise(false);
        } else {
          bonds(2);
          left_canonise(false);
        }
        max_truncation = 0.0;
      }
      ++step;
      bonds(1);
      left_canonise(false);
      bonds(0);
      left_canonise(step % interval == 0);
    }
  } else {
    for (int64_t n = 0; n < steps; ++n) {
      ++step;
      for (int k = 0; k < 10; ++k) {
        bonds(k);
        left_canonise(false);
      }
      bonds(10);
      left_canonise(step % interval == 0);
      max_truncation = 0.0;
    }
  }
}

}  // namespace cam
