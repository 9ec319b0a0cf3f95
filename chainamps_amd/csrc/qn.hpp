// QN charges and sector-decomposed tensor indices.
//
// Charge arithmetic restates State += / -= (states.hpp:176-205): a rule
// N > 1 is Z_N (values kept in [0,N)), N == 1 is trivial (always 0),
// N <= 0 is an integer ("momentum-like") U(1) charge.  An index's entries
// are grouped into sectors of equal charge; sectors are sorted by charge
// (lexicographic) and positions inside a sector keep their original order,
// so two indices with equal charge lists have identical sector tables.
#pragma once
#include <algorithm>
#include <array>
#include <cstdint>
#include <map>
#include <vector>

#include "common.hpp"

namespace cam {

constexpr int kMaxRules = 4;

struct Charge {
  std::array<int16_t, kMaxRules> q{};
  bool operator==(const Charge& o) const { return q == o.q; }
  bool operator!=(const Charge& o) const { return q != o.q; }
  bool operator<(const Charge& o) const { return q < o.q; }
};

struct Rules {
  std::vector<int16_t> N;
  int n() const { return static_cast<int>(N.size()); }
  bool operator==(const Rules& o) const { return N == o.N; }
  bool operator!=(const Rules& o) const { return N != o.N; }
  Charge zero() const { return Charge{}; }
  Charge add(const Charge& a, const Charge& b) const {
    Charge r;
    for (int i = 0; i < n(); ++i) {
      int v = int(a.q[i]) + int(b.q[i]);
      if (N[i] > 0) v %= N[i];
      r.q[i] = static_cast<int16_t>(v);
    }
    return r;
  }
  Charge sub(const Charge& a, const Charge& b) const {
    Charge r;
    for (int i = 0; i < n(); ++i) {
      int v = int(a.q[i]) - int(b.q[i]);
      if (N[i] > 0) {
        while (v < 0) v += N[i];
        v %= N[i];
      }
      r.q[i] = static_cast<int16_t>(v);
    }
    return r;
  }
  Charge neg(const Charge& a) const { return sub(zero(), a); }
  // constructor normalisation (states.cpp:13-27)
  Charge normalise(const int16_t* v) const {
    Charge r;
    for (int i = 0; i < n(); ++i) {
      int x = v[i];
      if (N[i] > 1) {
        CAM_REQUIRE(x >= 0, "Z_N charges can't be negative");
        x %= N[i];
      } else if (N[i] == 1) {
        x = 0;
      }
      r.q[i] = static_cast<int16_t>(x);
    }
    return r;
  }
};

// One tensor index with its sector table.
struct QIndex {
  bool inward = true;
  bool physical = false;
  int64_t size = 0;
  std::vector<Charge> charges;            // per position
  // sectors
  std::vector<Charge> sec_q;              // sorted
  std::vector<int64_t> sec_size;
  std::vector<std::vector<int64_t>> sec_pos;  // positions of each sector (ascending)
  std::vector<int32_t> pos_sec;           // position -> sector
  std::vector<int64_t> pos_off;           // position -> offset inside sector

  void build_sectors() {
    std::map<Charge, std::vector<int64_t>> m;
    for (int64_t p = 0; p < size; ++p) m[charges[p]].push_back(p);
    sec_q.clear(); sec_size.clear(); sec_pos.clear();
    pos_sec.assign(size, -1);
    pos_off.assign(size, -1);
    for (auto& kv : m) {
      int s = static_cast<int>(sec_q.size());
      sec_q.push_back(kv.first);
      sec_size.push_back(static_cast<int64_t>(kv.second.size()));
      for (size_t o = 0; o < kv.second.size(); ++o) {
        pos_sec[kv.second[o]] = s;
        pos_off[kv.second[o]] = static_cast<int64_t>(o);
      }
      sec_pos.push_back(std::move(kv.second));
    }
  }
  int nsec() const { return static_cast<int>(sec_q.size()); }
  int find_sector(const Charge& c) const {
    auto it = std::lower_bound(sec_q.begin(), sec_q.end(), c);
    if (it == sec_q.end() || *it != c) return -1;
    return static_cast<int>(it - sec_q.begin());
  }
  bool same_charges(const QIndex& o) const { return size == o.size && charges == o.charges; }
  QIndex flipped() const { QIndex r = *this; r.inward = !inward; return r; }
};

}  // namespace cam
