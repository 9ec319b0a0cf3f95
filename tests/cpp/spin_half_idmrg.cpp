// The reference's spin_half_example (examples/user_defined/spin_half_example,
// README.md:31-40) through the C++ layer that mirrors the reference API:
// H = -sum (sx sx + sy sy + sz sz) + 1e-5 sum sz on a spin-1/2 chain with
// Pauli operators, U(1) charge = 2 Sz, iDMRG growth with chi = 50 and the
// reference's growth-step eigensolver setting.  The README's growth
// energies -1, -3, -5, -7, -9 and entanglement entropies must come out.
// Also exercises contract / reorder / SVD / TwoVertexComponents::HED, the
// finite-chain TEBD on a product eigenstate (exact phase), and the iDMRG
// driver's Orthogonalise + iTwoVertexEnergy on the antiferromagnet (Bethe
// energy per vertex).
//
// Exit status 0 on success; prints one line per check.
#define CAM_AJAJ_THROW
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "chainamps_amd.hpp"

using namespace cam_ajaj;
typedef std::complex<double> C;

static int failures = 0;
static void expect(bool ok, const char* what, double got, double want) {
  std::printf("%s %s: got %.15g want %.15g\n", ok ? "ok  " : "FAIL", what, got, want);
  if (!ok) ++failures;
}

int main() {
  const std::vector<short> rules{0};   // one U(1) charge
  const StateArray phys{{1}, {-1}};    // up, down
  // sigma^+ (down -> up), sigma^- (up -> down), sigma^z, identity; d = 2
  auto op = [](int which, int r, int c) -> C {
    switch (which) {
      case 0: return (r == c) ? 1.0 : 0.0;                   // I
      case 1: return (r == 0 && c == 1) ? 1.0 : 0.0;         // s+
      case 2: return (r == 1 && c == 0) ? 1.0 : 0.0;         // s-
      default: return (r == c) ? (r == 0 ? 1.0 : -1.0) : 0.0;  // sz
    }
  };
  // lower-triangular MPO, w = 5: W[0,0] = I, W[4,4] = I, W[4,0] = hz sz,
  // bond terms J A_i B_{i+1}: W[4,j] = J A, W[j,0] = B with
  // (A, B, J) = (s+, s-, -2), (s-, s+, -2), (sz, sz, -1)
  const int d = 2, w = 5;
  const StateArray bq{{0}, {2}, {-2}, {0}, {0}};
  std::vector<C> Wd(static_cast<size_t>(d * w * d * w), 0.0);
  auto at = [&](int sp, int bl, int s, int br) -> C& { return Wd[sp + d * (bl + w * (s + d * br))]; };
  const int A[3] = {1, 2, 3}, B[3] = {2, 1, 3};
  const double J[3] = {-2.0, -2.0, -1.0};
  for (int sp = 0; sp < d; ++sp)
    for (int s = 0; s < d; ++s) {
      at(sp, 0, s, 0) = op(0, sp, s);
      at(sp, w - 1, s, w - 1) = op(0, sp, s);
      at(sp, w - 1, s, 0) = 1e-5 * op(3, sp, s);
      for (int c = 0; c < 3; ++c) {
        at(sp, w - 1, s, 1 + c) = J[c] * op(A[c], sp, s);
        at(sp, 1 + c, s, 0) = op(B[c], sp, s);
      }
    }
  const std::vector<MPXIndex> widx{MPXIndex(true, phys, true), MPXIndex(true, bq), MPXIndex(false, phys, true),
                                   MPXIndex(false, bq)};
  const MPX_matrix H = MPX_matrix::FromDense(rules, widx, 2, Wd);

  // growth energies and entropies of README.md:31-40
  const double energy[5] = {-1.0, -3.0, -5.0, -7.0, -9.0};
  const double entropy[5] = {0.6931471805599454, 0.8675632284814612, 1.018230153951394, 1.138073514962318,
                             1.235866276237509};
  FiniteDMRG dmrg(H, State{0}, /*converge=*/false);
  for (int step = 0; step < 5; ++step) {
    const iDMRGStep r = dmrg.grow(50);
    char what[64];
    std::snprintf(what, sizeof what, "iDMRG N=%lld energy", r.N);
    expect(std::fabs(r.Energy - energy[step]) < 1e-10, what, r.Energy, energy[step]);
    std::snprintf(what, sizeof what, "iDMRG N=%lld entropy", r.N);
    expect(std::fabs(r.Entropy - entropy[step]) < 1e-9, what, r.Entropy, entropy[step]);
  }
  // fDMRG_DRV -B 50 -F 2 (README.md:29-40): two finite sweeps of the ten
  // vertices; the midpoint rows (iterations 5-8) keep E = -9 and the
  // midpoint entropy 1.2358662762375
  {
    const std::vector<iDMRGStep> rows = dmrg.run(2, 50);
    expect(rows.size() == 4, "fDMRG midpoint rows", static_cast<double>(rows.size()), 4.0);
    for (const iDMRGStep& r : rows) {
      char what[80];
      std::snprintf(what, sizeof what, "fDMRG midpoint (%lld,%lld) energy", r.left, r.right);
      expect(r.left == 4 && r.right == 4 && std::fabs(r.Energy + 9.0) < 1e-10, what, r.Energy, -9.0);
      std::snprintf(what, sizeof what, "fDMRG midpoint (%lld,%lld) entropy", r.left, r.right);
      expect(std::fabs(r.Entropy - 1.235866276237508) < 1e-9, what, r.Entropy, 1.235866276237508);
    }
  }

  // the iDMRG driver's measurement step (iDMRG_DRV.cpp:64-82): Orthogonalise
  // the central unit cell and take iTwoVertexEnergy.  For the
  // antiferromagnet H = sum sigma.sigma (bond couplings +2, +2, +1 on
  // s+ s-, s- s+, sz sz, no field) the energy per vertex converges to
  // the Bethe value 1 - 4 ln 2 (4 (1/4 - ln 2) with Pauli operators).
  {
    std::vector<C> Wa(Wd.size(), 0.0);
    auto at2 = [&](int sp, int bl, int s, int br) -> C& { return Wa[sp + d * (bl + w * (s + d * br))]; };
    const double Ja[3] = {2.0, 2.0, 1.0};
    for (int sp = 0; sp < d; ++sp)
      for (int s = 0; s < d; ++s) {
        at2(sp, 0, s, 0) = op(0, sp, s);
        at2(sp, w - 1, s, w - 1) = op(0, sp, s);
        for (int c = 0; c < 3; ++c) {
          at2(sp, w - 1, s, 1 + c) = Ja[c] * op(A[c], sp, s);
          at2(sp, 1 + c, s, 0) = op(B[c], sp, s);
        }
      }
    const MPX_matrix Ha = MPX_matrix::FromDense(rules, widx, 2, Wa);
    iDMRG af(Ha, State{0}, /*converge=*/true);
    for (int step = 0; step < 40; ++step) af.grow(24);
    const UnitCell Ortho = Orthogonalise(af.CentralUnitCell());
    const double e = iTwoVertexEnergy(Ha, Ortho).real();
    const double bethe = 1.0 - 4.0 * std::log(2.0);
    expect(Ortho.size() == 2 && Ortho.Entropy() > 0.0, "Orthogonalise cell", Ortho.Entropy(), 0.0);
    expect(std::fabs(e - bethe) < 2e-4, "iTwoVertexEnergy antiferromagnet", e, bethe);
  }

  // contract / reorder / SVD on a two-site product: M = a (x) b over a
  // shared bond with known singular values
  const StateArray one{{0}};
  const StateArray bond{{0}, {0}, {0}};
  std::vector<C> a(6), b(6);
  for (int i = 0; i < 6; ++i) {
    a[i] = C(0.1 * (i + 1), 0.05 * i);
    b[i] = C(0.2 - 0.03 * i, -0.01 * i);
  }
  const MPX_matrix Ma = MPX_matrix::FromDense(rules, {MPXIndex(true, bond), MPXIndex(false, {{0}, {0}})}, 1, a);
  const MPX_matrix Mb = MPX_matrix::FromDense(rules, {MPXIndex(true, {{0}, {0}}), MPXIndex(false, bond)}, 1, b);
  const MPX_matrix AB = contract(Ma, false, Mb, false, MPXPair(1, 0));   // 3 x 3, rank 2
  const MPXDecomposition dec = reorder(AB, false, {0, 1}, 1).SVD();
  expect(dec.Values.size() == 2 || (dec.Values.size() == 3 && dec.Values[2] < 1e-14), "SVD rank of a (x) b",
         static_cast<double>(dec.Values.size()), 2.0);
  double fro2 = 0.0;
  std::vector<long long> p, i;
  std::vector<C> x;
  AB.to_csc(p, i, x);
  for (const C& v : x) fro2 += std::norm(v);
  double s2 = 0.0;
  for (double v : dec.Values) s2 += v * v;
  expect(std::fabs(s2 - fro2) < 1e-13 * fro2, "SVD sum s^2 = |AB|_F^2", s2, fro2);

  // store / load_MPX_matrix round trip of the MPO in the reference's format
  {
    const std::string path = "/tmp/cam_spin_half_H.MPO_matrix";
    H.store(path);
    const MPX_matrix H2 = load_MPX_matrix(path, rules, phys);
    std::vector<long long> p1, i1, p2, i2;
    std::vector<C> x1, x2;
    H.to_csc(p1, i1, x1);
    H2.to_csc(p2, i2, x2);
    std::remove(path.c_str());
    const bool same = p1 == p2 && i1 == i2 && x1 == x2 && H2.NumRowIndices() == H.NumRowIndices();
    expect(same, "store / load_MPX_matrix round trip (nnz)", static_cast<double>(x2.size()),
           static_cast<double>(x1.size()));
  }

  // finite-chain TEBD (TEBD::evolve) from the all-up product state of N = 4
  // sites: it is an eigenstate of every bond Hamiltonian (odd bonds carry the
  // field of both sites: -1 + 2e-5, even bonds -1), so each gate only adds a
  // phase and <psi(0)|psi(t)> = exp(-i E t) with E = -3 + 4e-5, at every
  // Trotter order, without truncation and with zero entanglement
  for (int order : {2, 4}) {
    std::vector<MPX_matrix> sites;
    for (int v = 1; v <= 4; ++v) {
      const std::vector<MPXIndex> sidx{MPXIndex(true, phys, true), MPXIndex(true, {{static_cast<short>(v - 1)}}),
                                       MPXIndex(false, {{static_cast<short>(v)}})};
      sites.push_back(MPX_matrix::FromDense(rules, sidx, 2, std::vector<C>{1.0, 0.0}));
    }
    const double dt = 0.1;
    TEBD tebd(H, sites, dt, order);
    tebd.evolve(5, 8, 0.0, 1);
    const std::vector<TEBDMeasurement> m = tebd.results();
    const double E = -3.0 + 4e-5;
    char what[96];
    std::snprintf(what, sizeof what, "TEBD order %d measurements", order);
    expect(m.size() == 5, what, static_cast<double>(m.size()), 5.0);
    if (m.empty()) continue;
    const TEBDMeasurement& last = m.back();
    const C want = std::exp(C(0.0, -E * last.Time));
    std::snprintf(what, sizeof what, "TEBD order %d overlap phase (re)", order);
    expect(std::abs(last.Overlap - want) < 1e-12, what, last.Overlap.real(), want.real());
    std::snprintf(what, sizeof what, "TEBD order %d entropy", order);
    expect(std::fabs(last.Entropy) < 1e-12 && last.Truncation < 1e-20, what, last.Entropy, 0.0);
  }

  std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
  return failures ? 1 : 0;
}
