// Host geometry restating include/sqsgen/core/structure.h for T = double and T = float
// (configuration<double> / configuration<float>, sqs.h:591-597).  Built with
// -ffp-contract=off: the reference is compiled for plain x86-64 (no FMA) and every
// shell assignment below follows its operation order.
#include "geometry.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <map>

namespace sqsg {

namespace {
// include/sqsgen/core/helpers/numeric.h:11-13
template <class T> inline bool is_close(T a, T b, T atol, T rtol) {
  return std::abs(a - b) <= (atol + rtol * std::abs(b));
}
}  // namespace

template <class T>
std::vector<T> distance_matrix(const LatticeT<T> &lat, const T *frac, uint32_t n) {
  const T *L = lat.m;
  // cart = frac * lattice (Eigen product, depth 3, summed in index order)
  std::vector<T> cart(3 * size_t(n));
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k)
      cart[3 * i + k] = frac[3 * i] * L[k] + frac[3 * i + 1] * L[3 + k] + frac[3 * i + 2] * L[6 + k];
  std::vector<T> d(size_t(n) * n, std::numeric_limits<T>::max());
  static const int axis[3] = {-1, 0, 1};
  for (uint32_t i = 0; i < n; ++i) {
    const T *p1 = &cart[3 * i];
    for (uint32_t j = i; j < n; ++j) {
      const T *p2 = &cart[3 * j];
      T best = d[size_t(i) * n + j];
      for (int u : axis)
        for (int v : axis)
          for (int w : axis) {
            T diff[3];
            for (int k = 0; k < 3; ++k) {
              const T t = T(u) * L[k] + T(v) * L[3 + k] + T(w) * L[6 + k];
              diff[k] = p1[k] - (t + p2[k]);
            }
            const T norm =
                std::abs(std::sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]));
            if (norm < best) best = norm;
          }
      d[size_t(i) * n + j] = best;
      d[size_t(j) * n + i] = best;
    }
  }
  return d;
}

template <class T>
std::vector<T> distances_naive(const std::vector<T> &dm, uint32_t, T atol, T rtol) {
  std::vector<T> u(dm);
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());  // sorted_vector keeps distinct values
  std::vector<T> out{T(0)};
  for (T x : u) {
    if (is_close(out.back(), x, atol, rtol))
      out.back() = static_cast<T>(0.5 * (x + out.back()));  // a double literal times T
    else
      out.push_back(x);
  }
  return out;
}

template <class T>
std::vector<T> distances_histogram(const std::vector<T> &dm, uint32_t, T bin_width,
                                   T peak_isolation) {
  std::vector<T> d;
  d.reserve(dm.size());
  for (T x : dm)
    if (x > 0.0 && !is_close<T>(x, 0.0, 1.0e-5, 1.0e-3)) d.push_back(x);
  if (d.empty()) throw std::invalid_argument("The structure contains no distinct sites");
  std::sort(d.begin(), d.end());
  const T lo = d.front(), hi = d.back();
  const size_t num_edges = static_cast<size_t>((hi - lo) / bin_width) + 2;
  if (num_edges < 10)
    throw std::invalid_argument(
        "Not enough edges to create a histogram, please increase the bin width");
  std::vector<T> edges(num_edges);
  for (size_t i = 0; i < num_edges; ++i) edges[i] = lo + static_cast<T>(i) * bin_width;
  // bin populations and maxima; bins are filled strictly in order
  const size_t nbins = num_edges - 1;
  std::vector<size_t> pop(nbins + 2, 0);
  std::vector<T> mx(nbins + 2, T(0));
  size_t index = 0, bin = 0;
  while (index < d.size() && bin < nbins) {
    const T lower = edges[bin], upper = edges[bin + 1], value = d[index];
    if (lower <= value && value < upper) {
      if (pop[bin] == 0 || value > mx[bin]) mx[bin] = value;
      ++pop[bin];
      ++index;
    } else {
      ++bin;
    }
  }
  auto population = [&](size_t i) -> size_t { return i <= nbins ? pop[i] : 0; };
  std::vector<T> shells;
  for (size_t i = 0; i <= num_edges; ++i) {
    const size_t prev = i == 0 ? 0 : population(i - 1);
    const size_t f = population(i), next = population(i + 1);
    // (1.0 - T) promotes to double, times T(count) (structure.h:246)
    const auto threshold = static_cast<size_t>((1.0 - peak_isolation) * static_cast<T>(f));
    if (threshold > prev && threshold > next) shells.push_back(mx[i]);
  }
  if (shells.empty()) throw std::invalid_argument("no isolated distance peaks found");
  if (shells.front() != 0.0 && !is_close<T>(shells.front(), 0.0, 1.0e-5, 1.0e-3))
    shells.insert(shells.begin(), T(0));
  return shells;
}

template <class T>
std::vector<uint32_t> shell_matrix(const std::vector<T> &dm, uint32_t n, const std::vector<T> &radii,
                                   T atol, T rtol) {
  auto find_shell = [&](T x) -> uint32_t {
    if (x < 0) throw std::out_of_range("Invalid distance matrix input");
    if (is_close<T>(x, 0.0, atol, rtol)) return 0;
    for (size_t r = 0; r + 1 < radii.size(); ++r) {
      const T lb = radii[r], ub = radii[r + 1];
      if ((is_close(x, lb, atol, rtol) || x > lb) && (is_close(x, ub, atol, rtol) || ub > x))
        return static_cast<uint32_t>(r + 1);
    }
    return static_cast<uint32_t>(radii.size());
  };
  std::vector<uint32_t> sm(size_t(n) * n, 0);
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j = i + 1; j < n; ++j) {
      const uint32_t s = find_shell(dm[size_t(i) * n + j]);
      sm[size_t(i) * n + j] = s;
      sm[size_t(j) * n + i] = s;
    }
  return sm;
}

template <class T>
std::vector<T> prefactors(const std::vector<uint32_t> &sm, uint32_t n,
                          const std::vector<uint32_t> &shells, const std::vector<uint8_t> &species,
                          uint32_t *num_species) {
  if (shells.empty()) throw std::out_of_range("no coordination shells selected");
  // neighbors: shell -> entries over the full N x N matrix, then entries / N stored
  // back into the size_t counter (truncation), structure.h:141-154
  std::map<uint32_t, size_t> neighbors;
  for (uint32_t s : sm) ++neighbors[s];
  for (auto &[shell, count] : neighbors)
    count = static_cast<size_t>(static_cast<T>(count) / static_cast<T>(n));
  std::map<uint8_t, size_t> hist;
  for (uint8_t z : species) ++hist[z];
  std::vector<uint8_t> zs;
  for (auto &[z, c] : hist) zs.push_back(z);
  const uint32_t K = static_cast<uint32_t>(zs.size());
  const size_t S = shells.size();
  std::vector<T> out(S * K * K, T(0));
  const T N = static_cast<T>(species.size());
  for (size_t s = 0; s < S; ++s) {
    auto it = neighbors.find(shells[s]);
    const size_t count = it == neighbors.end() ? 0 : it->second;
    if (count == 0) continue;  // zeros (the reference logs a warning)
    const T M = static_cast<T>(count);
    for (uint32_t a = 0; a < K; ++a) {
      const T xa = static_cast<T>(hist[zs[a]]) / N;
      for (uint32_t b = a; b < K; ++b) {
        const T xb = static_cast<T>(hist[zs[b]]) / N;
        const T p = T(1.0) / (M * xa * xb * N);
        out[(s * K + a) * K + b] = p;
        out[(s * K + b) * K + a] = p;
      }
    }
  }
  if (num_species) *num_species = K;
  return out;
}

template <class T>
void supercell(const LatticeT<T> &lat, const T *frac, const uint8_t *species, uint32_t n,
               uint32_t a, uint32_t b, uint32_t c, LatticeT<T> *lat_out, T *frac_out,
               uint8_t *species_out) {
  if (a == 0 || b == 0 || c == 0)
    throw std::invalid_argument("There must be at least one copy in each direction");
  // inverse of diag(a, b, c) the way Eigen's 3x3 cofactor inverse evaluates it
  const T sa = T(a), sb = T(b), sc = T(c);
  const T invdet = T(1) / ((sb * sc) * sa);
  const T inv[3] = {(sb * sc) * invdet, (sc * sa) * invdet, (sa * sb) * invdet};
  size_t site = 0;
  for (uint32_t x = 0; x < a; ++x)
    for (uint32_t y = 0; y < b; ++y)
      for (uint32_t z = 0; z < c; ++z) {
        const T t[3] = {T(x) * inv[0], T(y) * inv[1], T(z) * inv[2]};
        for (uint32_t q = 0; q < n; ++q, ++site) {
          for (int k = 0; k < 3; ++k) frac_out[3 * site + k] = t[k] + frac[3 * q + k] * inv[k];
          species_out[site] = species[q];
        }
      }
  const T s[3] = {sa, sb, sc};
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) lat_out->m[3 * r + k] = lat.m[3 * r + k] * s[k];
}

#define SQSG_GEOMETRY_INSTANCES(T)                                                          \
  template std::vector<T> distance_matrix<T>(const LatticeT<T> &, const T *, uint32_t);       \
  template std::vector<T> distances_naive<T>(const std::vector<T> &, uint32_t, T, T);         \
  template std::vector<T> distances_histogram<T>(const std::vector<T> &, uint32_t, T, T);     \
  template std::vector<uint32_t> shell_matrix<T>(const std::vector<T> &, uint32_t,            \
                                                 const std::vector<T> &, T, T);               \
  template std::vector<T> prefactors<T>(const std::vector<uint32_t> &, uint32_t,              \
                                        const std::vector<uint32_t> &,                        \
                                        const std::vector<uint8_t> &, uint32_t *);            \
  template void supercell<T>(const LatticeT<T> &, const T *, const uint8_t *, uint32_t,       \
                             uint32_t, uint32_t, uint32_t, LatticeT<T> *, T *, uint8_t *);
SQSG_GEOMETRY_INSTANCES(double)
SQSG_GEOMETRY_INSTANCES(float)
#undef SQSG_GEOMETRY_INSTANCES

}  // namespace sqsg
