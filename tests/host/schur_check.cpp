// Reads matrices "n re im re im ..." (column-major) from stdin, writes for
// each the Schur form T and vectors Z after ordering `count` eigenvalues of
// largest magnitude first: "n ok" then T then Z (re im pairs).
#include <cstdio>
#include <vector>

#include "../../chainamps_amd/csrc/schur.hpp"

int main() {
  int n, count;
  while (std::scanf("%d %d", &n, &count) == 2) {
    std::vector<cam::zcplx> A(static_cast<size_t>(n) * n), Z;
    for (auto& a : A) {
      double re, im;
      if (std::scanf("%lf %lf", &re, &im) != 2) return 1;
      a = {re, im};
    }
    const bool ok = cam::complex_schur(n, A, Z);
    if (ok) cam::schur_order_by_magnitude(n, A, Z, count);
    std::printf("%d %d\n", n, ok ? 1 : 0);
    for (const auto& a : A) std::printf("%.17g %.17g\n", a.real(), a.imag());
    for (const auto& z : Z) std::printf("%.17g %.17g\n", z.real(), z.imag());
  }
  return 0;
}
