// CPU baseline of the two-vertex H_eff matvec in the reference's own order
// (test and benchmark infrastructure, never part of the HIP product path).
//
// TwoVertexMPOMPSMultiply (DMRG_routines.cpp:898-974) forms the vector as a
// CSC matrix V with rows (a0, s1) and columns (s2, a2), multiplies the
// pre-contracted LeftPart = LeftBlock . W (rows (a0', s1', b), columns
// (a0, s1); contract3041 needs no reshape, MPX_matrix.cpp:696-698), shifts
// the row/column split of the product to rows (a0', s1') and columns
// (b, s2, a2) (ShiftNumRowIndices: the column-major order is unchanged),
// multiplies the pre-contracted RightPart = W . RightBlock (rows (b, s2, a2),
// columns (s2', a2'); contract203142 again needs no reshape) and gathers the
// allowed entries into the output vector (DumbExtractWithZeros).
//
// Both products are the reference's sparse_multiply (sparse_interface.cpp:
// 1892-2170) restated: a symbolic pass counts each answer column's nonzeros
// (Gustavson, one flag array per worker), the numeric pass scatters and
// gathers column by column, and the row order of the answer is restored
// either by sorting each column or, when nz(lhs) + nz(rhs) < nz(answer), by
// the swap-and-transpose route (C = (B^T A^T)^T).  Columns are split over
// worker threads in chunks, threading only when the reference's cutoff
// TBBNZ * ncols / nz(rhs) < ncols holds (TBBNZ = 1000,
// sparse_interface.cpp:18); std::thread with a shared chunk counter stands in
// for TBB's parallel_reduce over a blocked_range.
//
// The workload is the d-level Z2 coupled-chain model (levels of alternating
// parity, MPO bond w = 3 with channel parities 0, 1, 0: identity, J sigma,
// diag(E), sigma, identity) with seeded random complex environments of two
// parity sectors of chi/2; only the sparsity structure and sizes matter for
// the timing, and --check compares the result against a dense contraction.
//
// Usage: refmatvec --chi C --d D --threads T --reps R [--check]
// Prints one JSON line: timings per matvec, nonzeros, multiply-adds.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <thread>
#include <vector>

namespace {

using cplx = std::complex<double>;
constexpr int64_t kTbbNz = 1000;

struct Csc {
  int64_t m = 0, n = 0;
  std::vector<int64_t> p, i;
  std::vector<cplx> x;
  int64_t nz() const { return p.empty() ? 0 : p[n]; }
};

int g_threads = 1;

// parallel loop over [0, n) in chunks handed out by a shared counter; the
// body receives (begin, end, worker id)
void parallel_chunks(int64_t n, int64_t chunk, const std::function<void(int64_t, int64_t, int)>& body) {
  const int T = g_threads;
  if (T <= 1 || n <= chunk) {
    body(0, n, 0);
    return;
  }
  std::atomic<int64_t> next{0};
  auto work = [&](int id) {
    for (;;) {
      const int64_t b = next.fetch_add(chunk);
      if (b >= n) break;
      body(b, std::min(n, b + chunk), id);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
}

// column chunk: about 8 chunks per worker (TBB's auto partitioner splits a
// range into a few pieces per worker and steals the rest)
int64_t chunk_of(int64_t n) { return std::max<int64_t>(1, n / (8 * std::max(1, g_threads))); }

bool threaded(const Csc& B) { return g_threads > 1 && B.nz() > 0 && kTbbNz * B.n / B.nz() < B.n; }

Csc transpose(const Csc& A) {
  Csc T;
  T.m = A.n;
  T.n = A.m;
  T.p.assign(T.n + 1, 0);
  for (int64_t k = 0; k < A.nz(); ++k) T.p[A.i[k] + 1]++;
  for (int64_t r = 0; r < T.n; ++r) T.p[r + 1] += T.p[r];
  T.i.resize(A.nz());
  T.x.resize(A.nz());
  std::vector<int64_t> w(T.p.begin(), T.p.end() - 1);
  for (int64_t c = 0; c < A.n; ++c)
    for (int64_t k = A.p[c]; k < A.p[c + 1]; ++k) {
      const int64_t q = w[A.i[k]]++;
      T.i[q] = c;
      T.x[q] = A.x[k];
    }
  return T;
}

// symbolic pass: nonzeros of every column of A * B (cumulative)
std::vector<int64_t> count_nonzeros(const Csc& A, const Csc& B) {
  std::vector<int64_t> cp(B.n + 1, 0);
  auto body = [&](int64_t c0, int64_t c1, int) {
    std::vector<int64_t> flag(A.m, -1);
    for (int64_t col = c0; col < c1; ++col) {
      int64_t nzc = 0;
      for (int64_t rp = B.p[col]; rp < B.p[col + 1]; ++rp) {
        const int64_t j = B.i[rp];
        for (int64_t lp = A.p[j]; lp < A.p[j + 1]; ++lp)
          if (flag[A.i[lp]] != col) {
            flag[A.i[lp]] = col;
            ++nzc;
          }
      }
      cp[col + 1] = nzc;
    }
  };
  if (threaded(B)) parallel_chunks(B.n, chunk_of(B.n), body);
  else body(0, B.n, 0);
  for (int64_t c = 0; c < B.n; ++c) cp[c + 1] += cp[c];
  return cp;
}

// numeric pass into the column pointers of the symbolic pass
Csc numeric(const Csc& A, const Csc& B, std::vector<int64_t> cp, bool sort_rows, int64_t* macs) {
  Csc C;
  C.m = A.m;
  C.n = B.n;
  C.p = std::move(cp);
  C.i.resize(C.nz());
  C.x.resize(C.nz());
  std::vector<int64_t> mac_of(std::max(1, g_threads), 0);
  auto body = [&](int64_t c0, int64_t c1, int id) {
    std::vector<int64_t> flag(A.m, -1);
    std::vector<cplx> w(A.m);
    int64_t mac = 0;
    for (int64_t col = c0; col < c1; ++col) {
      int64_t q = C.p[col];
      for (int64_t rp = B.p[col]; rp < B.p[col + 1]; ++rp) {
        const int64_t j = B.i[rp];
        const cplx bx = B.x[rp];
        for (int64_t lp = A.p[j]; lp < A.p[j + 1]; ++lp) {
          const int64_t r = A.i[lp];
          if (flag[r] != col) {
            flag[r] = col;
            C.i[q++] = r;
            w[r] = A.x[lp] * bx;
          } else {
            w[r] += A.x[lp] * bx;
          }
        }
        mac += A.p[j + 1] - A.p[j];
      }
      if (sort_rows) std::sort(C.i.begin() + C.p[col], C.i.begin() + q);
      for (int64_t k = C.p[col]; k < q; ++k) C.x[k] = w[C.i[k]];
    }
    mac_of[id] += mac;
  };
  if (threaded(B)) parallel_chunks(B.n, chunk_of(B.n), body);
  else body(0, B.n, 0);
  for (int64_t v : mac_of) *macs += v;
  return C;
}

// sparse_multiply(lhs, rhs, NoSort = 0): answer with rows ordered in every column
Csc sparse_multiply(const Csc& A, const Csc& B, int64_t* macs) {
  std::vector<int64_t> cp = count_nonzeros(A, B);
  if (cp[B.n] == 0) {
    Csc C;
    C.m = A.m;
    C.n = B.n;
    C.p.assign(B.n + 1, 0);
    return C;
  }
  if (A.nz() + B.nz() < cp[B.n]) {
    // swap and transpose: (B^T A^T) computed unsorted, then one transpose
    // orders the rows
    const Csc At = transpose(A), Bt = transpose(B);
    Csc Ct = numeric(Bt, At, count_nonzeros(Bt, At), false, macs);
    return transpose(Ct);
  }
  return numeric(A, B, std::move(cp), true, macs);
}

struct Model {
  int64_t chi, d, w = 3;
  std::vector<cplx> L, R, W;   // L[a', b, a], R[a', b, a], W[s', bl, s, br] dense
  int pa(int64_t a) const { return a >= chi / 2 ? 1 : 0; }
  int ps(int64_t s) const { return static_cast<int>(s & 1); }
  int pb(int64_t b) const { return b == 1 ? 1 : 0; }
  cplx& l(int64_t a1, int64_t b, int64_t a) { return L[(a1 * w + b) * chi + a]; }
  cplx& r(int64_t a1, int64_t b, int64_t a) { return R[(a1 * w + b) * chi + a]; }
  cplx& wt(int64_t s1, int64_t bl, int64_t s, int64_t br) { return W[((s1 * w + bl) * d + s) * w + br]; }
  bool psi_allowed(int64_t a0, int64_t s1, int64_t s2, int64_t a2) const {
    return (pa(a0) ^ ps(s1) ^ ps(s2) ^ pa(a2)) == 0;
  }
};

Model make_model(int64_t chi, int64_t d, uint64_t seed) {
  Model M;
  M.chi = chi;
  M.d = d;
  const int64_t w = M.w;
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd;
  M.L.assign(chi * w * chi, 0.0);
  M.R.assign(chi * w * chi, 0.0);
  for (int64_t a1 = 0; a1 < chi; ++a1)
    for (int64_t b = 0; b < w; ++b)
      for (int64_t a = 0; a < chi; ++a)
        if ((M.pa(a1) ^ M.pb(b) ^ M.pa(a)) == 0) {
          M.l(a1, b, a) = cplx(nd(rng), nd(rng));
          M.r(a1, b, a) = cplx(nd(rng), nd(rng));
        }
  std::vector<double> E(d), sig(d * d, 0.0);
  std::uniform_real_distribution<double> ud(0.0, 8.0);
  for (auto& e : E) e = ud(rng);
  std::sort(E.begin(), E.end());
  for (int64_t i = 0; i < d; ++i)
    for (int64_t j = 0; j <= i; ++j)
      if (M.ps(i) != M.ps(j)) sig[i * d + j] = sig[j * d + i] = nd(rng);
  const double J = 0.1;
  M.W.assign(d * w * d * w, 0.0);
  for (int64_t s1 = 0; s1 < d; ++s1)
    for (int64_t s = 0; s < d; ++s) {
      if (s1 == s) {
        M.wt(s1, 0, s, 0) = 1.0;
        M.wt(s1, 2, s, 2) = 1.0;
        M.wt(s1, 2, s, 0) = E[s];
      }
      M.wt(s1, 1, s, 0) = J * sig[s1 * d + s];
      M.wt(s1, 2, s, 1) = sig[s1 * d + s];
    }
  return M;
}

// LeftPart: rows (a0', s1', b), columns (a0, s1)
Csc left_part(Model& M) {
  const int64_t chi = M.chi, d = M.d, w = M.w;
  Csc A;
  A.m = chi * d * w;
  A.n = chi * d;
  A.p.assign(A.n + 1, 0);
  for (int64_t s1 = 0; s1 < d; ++s1)
    for (int64_t a0 = 0; a0 < chi; ++a0) {
      for (int64_t b = 0; b < w; ++b)
        for (int64_t s1p = 0; s1p < d; ++s1p)
          for (int64_t a0p = 0; a0p < chi; ++a0p) {
            cplx v = 0.0;
            bool any = false;
            for (int64_t bl = 0; bl < w; ++bl) {
              const cplx wv = M.wt(s1p, bl, s1, b);
              const cplx lv = M.l(a0p, bl, a0);
              if (wv != 0.0 && lv != 0.0) {
                v += lv * wv;
                any = true;
              }
            }
            if (any) {
              A.i.push_back(a0p + chi * s1p + chi * d * b);
              A.x.push_back(v);
            }
          }
      A.p[a0 + chi * s1 + 1] = static_cast<int64_t>(A.i.size());
    }
  return A;
}

// RightPart: rows (b, s2, a2), columns (s2', a2')
Csc right_part(Model& M) {
  const int64_t chi = M.chi, d = M.d, w = M.w;
  Csc B;
  B.m = w * d * chi;
  B.n = d * chi;
  B.p.assign(B.n + 1, 0);
  for (int64_t a2p = 0; a2p < chi; ++a2p)
    for (int64_t s2p = 0; s2p < d; ++s2p) {
      for (int64_t a2 = 0; a2 < chi; ++a2)
        for (int64_t s2 = 0; s2 < d; ++s2)
          for (int64_t b = 0; b < w; ++b) {
            cplx v = 0.0;
            bool any = false;
            for (int64_t br = 0; br < w; ++br) {
              const cplx wv = M.wt(s2p, b, s2, br);
              const cplx rv = M.r(a2p, br, a2);
              if (wv != 0.0 && rv != 0.0) {
                v += wv * rv;
                any = true;
              }
            }
            if (any) {
              B.i.push_back(b + w * s2 + w * d * a2);
              B.x.push_back(v);
            }
          }
      B.p[s2p + d * a2p + 1] = static_cast<int64_t>(B.i.size());
    }
  return B;
}

struct Allowed {
  std::vector<int64_t> row, col;   // (a0 + chi s1, s2 + d a2), column-major order
};

Allowed allowed_entries(const Model& M) {
  Allowed al;
  for (int64_t a2 = 0; a2 < M.chi; ++a2)
    for (int64_t s2 = 0; s2 < M.d; ++s2)
      for (int64_t s1 = 0; s1 < M.d; ++s1)
        for (int64_t a0 = 0; a0 < M.chi; ++a0)
          if (M.psi_allowed(a0, s1, s2, a2)) {
            al.row.push_back(a0 + M.chi * s1);
            al.col.push_back(s2 + M.d * a2);
          }
  return al;
}

struct Timing {
  double form = 0, mult1 = 0, shift = 0, mult2 = 0, extract = 0;
  int64_t macs = 0, nzM = 0, nzOut = 0;
};

double secs(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

void matvec(const Model& M, const Csc& LP, const Csc& RP, const Allowed& al, const cplx* in, cplx* out,
            Timing& t) {
  const int64_t chi = M.chi, d = M.d, w = M.w;
  auto t0 = std::chrono::steady_clock::now();
  // the vector as a CSC matrix: allowed entries are already column-major
  Csc V;
  V.m = chi * d;
  V.n = d * chi;
  V.p.assign(V.n + 1, 0);
  const int64_t n = static_cast<int64_t>(al.row.size());
  for (int64_t k = 0; k < n; ++k) V.p[al.col[k] + 1]++;
  for (int64_t c = 0; c < V.n; ++c) V.p[c + 1] += V.p[c];
  V.i = al.row;
  V.x.assign(in, in + n);
  auto t1 = std::chrono::steady_clock::now();
  Csc P = sparse_multiply(LP, V, &t.macs);
  auto t2 = std::chrono::steady_clock::now();
  // ShiftNumRowIndices(2): rows (a0', s1'), columns (b, s2, a2); each old
  // column splits into w new ones, rows already ordered
  Csc S;
  S.m = chi * d;
  S.n = w * V.n;
  S.p.assign(S.n + 1, 0);
  S.i.resize(P.nz());
  S.x = std::move(P.x);
  for (int64_t c = 0; c < P.n; ++c)
    for (int64_t k = P.p[c]; k < P.p[c + 1]; ++k) {
      const int64_t b = P.i[k] / (chi * d);
      S.i[k] = P.i[k] - b * chi * d;
      S.p[c * w + b + 1]++;
    }
  for (int64_t c = 0; c < S.n; ++c) S.p[c + 1] += S.p[c];
  t.nzM = S.nz();
  auto t3 = std::chrono::steady_clock::now();
  Csc O = sparse_multiply(S, RP, &t.macs);
  t.nzOut = O.nz();
  auto t4 = std::chrono::steady_clock::now();
  // DumbExtractWithZeros: every allowed entry, zero where the answer has none
  for (int64_t k = 0; k < n; ++k) {
    const int64_t c = al.col[k];
    const auto b = O.i.begin() + O.p[c], e = O.i.begin() + O.p[c + 1];
    const auto it = std::lower_bound(b, e, al.row[k]);
    out[k] = (it != e && *it == al.row[k]) ? O.x[it - O.i.begin()] : cplx(0.0);
  }
  auto t5 = std::chrono::steady_clock::now();
  t.form += secs(t0, t1);
  t.mult1 += secs(t1, t2);
  t.shift += secs(t2, t3);
  t.mult2 += secs(t3, t4);
  t.extract += secs(t4, t5);
}

// dense contraction of the same model (check mode, small sizes)
double check(Model& M, const Allowed& al, const std::vector<cplx>& in, const std::vector<cplx>& out) {
  const int64_t chi = M.chi, d = M.d, w = M.w;
  std::vector<cplx> psi(chi * d * d * chi, 0.0), ref(chi * d * d * chi, 0.0);
  auto at = [&](int64_t a0, int64_t s1, int64_t s2, int64_t a2) { return a0 + chi * (s1 + d * (s2 + d * a2)); };
  for (size_t k = 0; k < al.row.size(); ++k) {
    const int64_t a0 = al.row[k] % chi, s1 = al.row[k] / chi, s2 = al.col[k] % d, a2 = al.col[k] / d;
    psi[at(a0, s1, s2, a2)] = in[k];
  }
  for (int64_t a0p = 0; a0p < chi; ++a0p)
    for (int64_t s1p = 0; s1p < d; ++s1p)
      for (int64_t s2p = 0; s2p < d; ++s2p)
        for (int64_t a2p = 0; a2p < chi; ++a2p) {
          cplx acc = 0.0;
          for (int64_t a0 = 0; a0 < chi; ++a0)
            for (int64_t s1 = 0; s1 < d; ++s1)
              for (int64_t s2 = 0; s2 < d; ++s2)
                for (int64_t a2 = 0; a2 < chi; ++a2) {
                  const cplx v = psi[at(a0, s1, s2, a2)];
                  if (v == 0.0) continue;
                  cplx h = 0.0;
                  for (int64_t bl = 0; bl < w; ++bl)
                    for (int64_t b = 0; b < w; ++b)
                      for (int64_t br = 0; br < w; ++br)
                        h += M.l(a0p, bl, a0) * M.wt(s1p, bl, s1, b) * M.wt(s2p, b, s2, br) * M.r(a2p, br, a2);
                  acc += h * v;
                }
          ref[at(a0p, s1p, s2p, a2p)] = acc;
        }
  double num = 0, den = 0;
  for (size_t k = 0; k < al.row.size(); ++k) {
    const int64_t a0 = al.row[k] % chi, s1 = al.row[k] / chi, s2 = al.col[k] % d, a2 = al.col[k] / d;
    const cplx r = ref[at(a0, s1, s2, a2)];
    num += std::norm(out[k] - r);
    den += std::norm(r);
  }
  // entries outside the allowed set vanish by the Z2 rule
  double outside = 0;
  for (int64_t a0 = 0; a0 < chi; ++a0)
    for (int64_t s1 = 0; s1 < d; ++s1)
      for (int64_t s2 = 0; s2 < d; ++s2)
        for (int64_t a2 = 0; a2 < chi; ++a2)
          if (!M.psi_allowed(a0, s1, s2, a2)) outside += std::norm(ref[at(a0, s1, s2, a2)]);
  return std::sqrt((num + outside) / den);
}

}  // namespace

int main(int argc, char** argv) {
  int64_t chi = 16, d = 64;
  int reps = 1;
  bool do_check = false;
  uint64_t seed = 0;
  for (int a = 1; a < argc; ++a) {
    const std::string s = argv[a];
    auto next = [&]() -> const char* {
      if (a + 1 >= argc) {
        std::fprintf(stderr, "missing value after %s\n", s.c_str());
        std::exit(2);
      }
      return argv[++a];
    };
    if (s == "--chi") chi = std::atoll(next());
    else if (s == "--d") d = std::atoll(next());
    else if (s == "--threads") g_threads = std::atoi(next());
    else if (s == "--reps") reps = std::atoi(next());
    else if (s == "--seed") seed = std::strtoull(next(), nullptr, 10);
    else if (s == "--check") do_check = true;
    else {
      std::fprintf(stderr, "unknown argument %s\n", s.c_str());
      return 2;
    }
  }
  if (chi < 2 || chi % 2 || d < 2 || reps < 1 || g_threads < 1) {
    std::fprintf(stderr, "need even chi >= 2, d >= 2, reps >= 1, threads >= 1\n");
    return 2;
  }
  auto s0 = std::chrono::steady_clock::now();
  Model M = make_model(chi, d, seed);
  const Csc LP = left_part(M);
  const Csc RP = right_part(M);
  const Allowed al = allowed_entries(M);
  const int64_t n = static_cast<int64_t>(al.row.size());
  std::vector<cplx> in(n), out(n);
  std::mt19937_64 rng(seed + 1);
  std::normal_distribution<double> nd;
  for (auto& v : in) v = cplx(nd(rng), nd(rng));
  auto s1 = std::chrono::steady_clock::now();
  Timing t;
  auto m0 = std::chrono::steady_clock::now();
  for (int r = 0; r < reps; ++r) matvec(M, LP, RP, al, in.data(), out.data(), t);
  auto m1 = std::chrono::steady_clock::now();
  const double per = secs(m0, m1) / reps;
  std::printf("{\"chi\": %lld, \"d\": %lld, \"threads\": %d, \"reps\": %d, \"n\": %lld, \"s_per_matvec\": %.6g, "
              "\"stages_s\": {\"form_vector\": %.6g, \"multiply_left\": %.6g, \"shift\": %.6g, "
              "\"multiply_right\": %.6g, \"extract\": %.6g}, \"complex_macs\": %lld, \"nz_left_part\": %lld, "
              "\"nz_right_part\": %lld, \"nz_intermediate\": %lld, \"nz_out\": %lld, \"setup_s\": %.4g",
              static_cast<long long>(chi), static_cast<long long>(d), g_threads, reps, static_cast<long long>(n),
              per, t.form / reps, t.mult1 / reps, t.shift / reps, t.mult2 / reps, t.extract / reps,
              static_cast<long long>(t.macs / reps), static_cast<long long>(LP.nz()),
              static_cast<long long>(RP.nz()), static_cast<long long>(t.nzM), static_cast<long long>(t.nzOut),
              secs(s0, s1));
  if (do_check) std::printf(", \"check_rel_err\": %.3e", check(M, al, in, out));
  std::printf("}\n");
  return 0;
}
