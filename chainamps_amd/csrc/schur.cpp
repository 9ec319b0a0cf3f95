// Complex Schur decomposition and reordering (host, small matrices).
#include "schur.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace cam {

namespace {

// [c s; -conj(s) c] [x; y] = [r; 0] with c real (LAPACK zlartg's rotation)
void givens(zcplx x, zcplx y, double& c, zcplx& s) {
  const double ax = std::abs(x), ay = std::abs(y);
  if (ay == 0.0) {
    c = 1.0;
    s = 0.0;
    return;
  }
  if (ax == 0.0) {
    c = 0.0;
    s = 1.0;
    return;
  }
  const double r = std::hypot(ax, ay);
  c = ax / r;
  s = (x / ax) * std::conj(y) / r;
}

struct Mat {
  int n;
  std::vector<zcplx>& a;
  zcplx& operator()(int i, int j) { return a[static_cast<size_t>(j) * n + i]; }
};

// rows k, k+1 <- G rows, columns j0..n-1
void rot_rows(Mat& M, int k, int j0, double c, zcplx s) {
  for (int j = j0; j < M.n; ++j) {
    const zcplx t1 = M(k, j), t2 = M(k + 1, j);
    M(k, j) = c * t1 + s * t2;
    M(k + 1, j) = -std::conj(s) * t1 + c * t2;
  }
}

// columns k, k+1 <- columns G^H, rows 0..i1
void rot_cols(Mat& M, int k, int i1, double c, zcplx s) {
  for (int i = 0; i <= i1; ++i) {
    const zcplx t1 = M(i, k), t2 = M(i, k + 1);
    M(i, k) = c * t1 + std::conj(s) * t2;
    M(i, k + 1) = -s * t1 + c * t2;
  }
}

}  // namespace

bool complex_schur(int n, std::vector<zcplx>& A, std::vector<zcplx>& Z) {
  Z.assign(static_cast<size_t>(n) * n, 0.0);
  Mat H{n, A}, Q{n, Z};
  for (int i = 0; i < n; ++i) Q(i, i) = 1.0;
  // Hessenberg reduction: H_k = I - 2 v v^H / |v|^2 zeroes column k below k+1
  std::vector<zcplx> v(n);
  for (int k = 0; k + 2 < n; ++k) {
    double xn = 0.0;
    for (int i = k + 1; i < n; ++i) xn += std::norm(H(i, k));
    xn = std::sqrt(xn);
    if (xn == 0.0) continue;
    const zcplx x0 = H(k + 1, k);
    const zcplx alpha = -(std::abs(x0) > 0.0 ? x0 / std::abs(x0) : zcplx(1.0)) * xn;
    const int m = n - k - 1;
    v[0] = x0 - alpha;
    double vn = std::norm(v[0]);
    for (int i = 1; i < m; ++i) {
      v[i] = H(k + 1 + i, k);
      vn += std::norm(v[i]);
    }
    if (vn == 0.0) continue;
    const double f = 2.0 / vn;
    for (int j = k; j < n; ++j) {
      zcplx s = 0.0;
      for (int i = 0; i < m; ++i) s += std::conj(v[i]) * H(k + 1 + i, j);
      s *= f;
      for (int i = 0; i < m; ++i) H(k + 1 + i, j) -= v[i] * s;
    }
    for (Mat* M : {&H, &Q})
      for (int i = 0; i < n; ++i) {
        zcplx s = 0.0;
        for (int l = 0; l < m; ++l) s += (*M)(i, k + 1 + l) * v[l];
        s *= f;
        for (int l = 0; l < m; ++l) (*M)(i, k + 1 + l) -= s * std::conj(v[l]);
      }
    H(k + 1, k) = alpha;
    for (int i = k + 2; i < n; ++i) H(i, k) = 0.0;
  }
  // QR sweeps on the active block [l, hi], updating the whole T and Z
  const double eps = std::numeric_limits<double>::epsilon();
  int hi = n - 1, iter = 0, total = 0;
  while (hi > 0) {
    int l = hi;
    for (; l > 0; --l) {
      double s = std::abs(H(l - 1, l - 1)) + std::abs(H(l, l));
      if (s == 0.0) s = std::numeric_limits<double>::min();
      if (std::abs(H(l, l - 1)) <= eps * s) {
        H(l, l - 1) = 0.0;
        break;
      }
    }
    if (l == hi) {
      --hi;
      iter = 0;
      continue;
    }
    if (++total > 30 * n) return false;
    ++iter;
    zcplx mu;
    if (iter % 11 == 10) {
      mu = H(hi, hi) + 0.75 * std::abs(H(hi, hi - 1));
    } else {
      const zcplx a = H(hi - 1, hi - 1), b = H(hi - 1, hi), c = H(hi, hi - 1), d = H(hi, hi);
      const zcplx half = 0.5 * (a + d);
      const zcplx disc = std::sqrt(0.25 * (a - d) * (a - d) + b * c);
      const zcplx m1 = half + disc, m2 = half - disc;
      mu = std::abs(m1 - d) < std::abs(m2 - d) ? m1 : m2;
    }
    zcplx x = H(l, l) - mu, y = H(l + 1, l);
    for (int k = l; k < hi; ++k) {
      double c;
      zcplx s;
      givens(x, y, c, s);
      rot_rows(H, k, k > l ? k - 1 : l, c, s);
      rot_cols(H, k, std::min(k + 2, hi), c, s);
      rot_cols(Q, k, n - 1, c, s);
      if (k > l) H(k + 1, k - 1) = 0.0;
      if (k + 1 < hi) {
        x = H(k + 1, k);
        y = H(k + 2, k);
      }
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) H(i, j) = 0.0;
  return true;
}

void schur_swap(int n, std::vector<zcplx>& T, std::vector<zcplx>& Z, int k) {
  Mat M{n, T}, Q{n, Z};
  const zcplx a = M(k, k), b = M(k + 1, k + 1);
  if (a == b) return;
  double c;
  zcplx s;
  givens(M(k, k + 1), b - a, c, s);
  rot_rows(M, k, k, c, s);
  rot_cols(M, k, k + 1, c, s);
  rot_cols(Q, k, n - 1, c, s);
  M(k + 1, k) = 0.0;
  M(k, k) = b;
  M(k + 1, k + 1) = a;
}

void schur_order_by_magnitude(int n, std::vector<zcplx>& T, std::vector<zcplx>& Z, int count) {
  Mat M{n, T};
  for (int p = 0; p < std::min(count, n); ++p) {
    int best = p;
    for (int j = p + 1; j < n; ++j)
      if (std::abs(M(j, j)) > std::abs(M(best, best))) best = j;
    for (int j = best - 1; j >= p; --j) schur_swap(n, T, Z, j);
  }
}

}  // namespace cam
