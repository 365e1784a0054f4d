// Multiset-permutation arithmetic on the host (include/sqsgen/core/permutation.h:17-102).
// rank_t is boost::multiprecision::cpp_int in the reference (include/sqsgen/types.h:20);
// reported ranks of whole configurations can exceed 128 bits, so they use BigUint.
// Systematic enumeration is capped at 2^64 - 1 permutations (io/config/combined.h:90-94),
// so chunk-start unranking uses 128-bit arithmetic.
#pragma once

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace sqsg {

using u128 = unsigned __int128;

// Non-negative arbitrary-precision integer, base 2^32 little-endian limbs.
class BigUint {
 public:
  BigUint(uint64_t v = 0) {
    while (v) {
      limbs_.push_back(static_cast<uint32_t>(v));
      v >>= 32;
    }
  }
  static BigUint from_string(const std::string &s) {
    BigUint r;
    if (s.empty()) throw std::invalid_argument("empty rank string");
    for (char ch : s) {
      if (ch < '0' || ch > '9') throw std::invalid_argument("rank must be a decimal integer");
      r.mul_small(10);
      r.add(BigUint(static_cast<uint64_t>(ch - '0')));
    }
    return r;
  }
  bool is_zero() const { return limbs_.empty(); }
  void mul_small(uint32_t m) {
    uint64_t carry = 0;
    for (auto &l : limbs_) {
      const uint64_t t = uint64_t(l) * m + carry;
      l = static_cast<uint32_t>(t);
      carry = t >> 32;
    }
    if (carry) limbs_.push_back(static_cast<uint32_t>(carry));
    trim();
  }
  uint32_t div_small(uint32_t d) {  // returns the remainder
    uint64_t rem = 0;
    for (size_t i = limbs_.size(); i-- > 0;) {
      const uint64_t cur = (rem << 32) | limbs_[i];
      limbs_[i] = static_cast<uint32_t>(cur / d);
      rem = cur % d;
    }
    trim();
    return static_cast<uint32_t>(rem);
  }
  void add(const BigUint &o) {
    uint64_t carry = 0;
    if (limbs_.size() < o.limbs_.size()) limbs_.resize(o.limbs_.size(), 0);
    for (size_t i = 0; i < limbs_.size(); ++i) {
      const uint64_t t = uint64_t(limbs_[i]) + (i < o.limbs_.size() ? o.limbs_[i] : 0) + carry;
      limbs_[i] = static_cast<uint32_t>(t);
      carry = t >> 32;
    }
    if (carry) limbs_.push_back(static_cast<uint32_t>(carry));
  }
  bool fits_u64() const { return limbs_.size() <= 2; }
  bool fits_u128() const { return limbs_.size() <= 4; }
  u128 to_u128() const {
    if (!fits_u128()) throw std::out_of_range("value exceeds 128 bits");
    u128 v = 0;
    for (size_t i = limbs_.size(); i-- > 0;) v = (v << 32) | limbs_[i];
    return v;
  }
  std::string str() const {
    if (limbs_.empty()) return "0";
    BigUint t(*this);
    std::string s;
    while (!t.is_zero()) {
      s.push_back(static_cast<char>('0' + t.div_small(10)));
    }
    std::reverse(s.begin(), s.end());
    return s;
  }

 private:
  void trim() {
    while (!limbs_.empty() && limbs_.back() == 0) limbs_.pop_back();
  }
  std::vector<uint32_t> limbs_;
};

inline std::string u128_str(u128 v) {
  if (v == 0) return "0";
  std::string s;
  while (v) {
    s.push_back(static_cast<char>('0' + static_cast<int>(v % 10)));
    v /= 10;
  }
  std::reverse(s.begin(), s.end());
  return s;
}

// permutation.h:17-32, N!/prod(n_k!) built as a product of binomials (exact at every step)
inline BigUint num_permutations(const uint8_t *conf, size_t n) {
  size_t hist[256] = {0};
  for (size_t i = 0; i < n; ++i) ++hist[conf[i]];
  BigUint total(1);
  uint32_t placed = 0;
  for (int s = 0; s < 256; ++s)
    for (size_t k = 1; k <= hist[s]; ++k) {
      ++placed;
      total.mul_small(placed);
      total.div_small(static_cast<uint32_t>(k));
    }
  return total;
}

// permutation.h:34-50: 1-based lexicographic rank; species must be dense 0..K-1
inline BigUint rank_permutation(const uint8_t *conf, size_t n) {
  bool seen[256] = {false};
  uint32_t num_species = 0;
  for (size_t i = 0; i < n; ++i)
    if (!seen[conf[i]]) {
      seen[conf[i]] = true;
      ++num_species;
    }
  std::vector<uint32_t> hist(256, 0);
  BigUint rank(1), suffix(1);
  for (size_t i = 0; i < n; ++i) {
    const uint8_t x = conf[(n - 1) - i];
    ++hist[x];
    for (uint32_t j = 0; j < num_species && j < x; ++j) {
      if (hist[j] == 0) continue;
      BigUint t(suffix);
      t.mul_small(hist[j]);
      t.div_small(hist[x]);
      rank.add(t);
    }
    suffix.mul_small(static_cast<uint32_t>(i + 1));
    suffix.div_small(hist[x]);
  }
  return rank;
}

// permutation.h:52-78 on 128-bit ranks; writes dense indices of the present values
inline void unrank_permutation(uint8_t *conf, size_t n, u128 rank) {
  size_t counts[256] = {0};
  for (size_t i = 0; i < n; ++i) ++counts[conf[i]];
  std::vector<uint64_t> hist;
  for (int s = 0; s < 256; ++s)
    if (counts[s]) hist.push_back(counts[s]);
  const BigUint tb = num_permutations(conf, n);
  if (!tb.fits_u128()) throw std::out_of_range("too many permutations for 128-bit ranks");
  u128 total = tb.to_u128();
  if (rank > total)
    throw std::out_of_range("The rank is larger than the total number of permutations");
  size_t k = 0;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < hist.size(); ++j) {
      const u128 suffix_count = total * hist[j] / (n - i);
      if (rank <= suffix_count) {
        conf[k++] = static_cast<uint8_t>(j);
        total = suffix_count;
        --hist[j];
        break;
      }
      rank -= suffix_count;
    }
}

}  // namespace sqsg
