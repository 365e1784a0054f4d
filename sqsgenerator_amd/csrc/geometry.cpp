// Host geometry restating include/sqsgen/core/structure.h.  Built with
// -ffp-contract=off: the reference is compiled for plain x86-64 (no FMA) and every
// shell assignment below follows its operation order.
#include "geometry.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <map>

namespace sqsg {

namespace {
// include/sqsgen/core/helpers/numeric.h:11-13
inline bool is_close(double a, double b, double atol, double rtol) {
  return std::abs(a - b) <= (atol + rtol * std::abs(b));
}
}  // namespace

std::vector<double> distance_matrix(const Lattice &lat, const double *frac, uint32_t n) {
  const double *L = lat.m;
  // cart = frac * lattice (Eigen product, depth 3, summed in index order)
  std::vector<double> cart(3 * size_t(n));
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k)
      cart[3 * i + k] = frac[3 * i] * L[k] + frac[3 * i + 1] * L[3 + k] + frac[3 * i + 2] * L[6 + k];
  std::vector<double> d(size_t(n) * n, DBL_MAX);
  static const int axis[3] = {-1, 0, 1};
  for (uint32_t i = 0; i < n; ++i) {
    const double *p1 = &cart[3 * i];
    for (uint32_t j = i; j < n; ++j) {
      const double *p2 = &cart[3 * j];
      double best = d[size_t(i) * n + j];
      for (int u : axis)
        for (int v : axis)
          for (int w : axis) {
            double diff[3];
            for (int k = 0; k < 3; ++k) {
              const double t = double(u) * L[k] + double(v) * L[3 + k] + double(w) * L[6 + k];
              diff[k] = p1[k] - (t + p2[k]);
            }
            const double norm =
                std::abs(std::sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]));
            if (norm < best) best = norm;
          }
      d[size_t(i) * n + j] = best;
      d[size_t(j) * n + i] = best;
    }
  }
  return d;
}

std::vector<double> distances_naive(const std::vector<double> &dm, uint32_t, double atol,
                                    double rtol) {
  std::vector<double> u(dm);
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());  // sorted_vector keeps distinct values
  std::vector<double> out{0.0};
  for (double x : u) {
    if (is_close(out.back(), x, atol, rtol))
      out.back() = 0.5 * (x + out.back());
    else
      out.push_back(x);
  }
  return out;
}

std::vector<double> distances_histogram(const std::vector<double> &dm, uint32_t, double bin_width,
                                        double peak_isolation) {
  std::vector<double> d;
  d.reserve(dm.size());
  for (double x : dm)
    if (x > 0.0 && !is_close(x, 0.0, 1.0e-5, 1.0e-3)) d.push_back(x);
  if (d.empty()) throw std::invalid_argument("The structure contains no distinct sites");
  std::sort(d.begin(), d.end());
  const double lo = d.front(), hi = d.back();
  const size_t num_edges = static_cast<size_t>((hi - lo) / bin_width) + 2;
  if (num_edges < 10)
    throw std::invalid_argument(
        "Not enough edges to create a histogram, please increase the bin width");
  std::vector<double> edges(num_edges);
  for (size_t i = 0; i < num_edges; ++i) edges[i] = lo + static_cast<double>(i) * bin_width;
  // bin populations and maxima; bins are filled strictly in order
  const size_t nbins = num_edges - 1;
  std::vector<size_t> pop(nbins + 2, 0);
  std::vector<double> mx(nbins + 2, 0.0);
  size_t index = 0, bin = 0;
  while (index < d.size() && bin < nbins) {
    const double lower = edges[bin], upper = edges[bin + 1], value = d[index];
    if (lower <= value && value < upper) {
      if (pop[bin] == 0 || value > mx[bin]) mx[bin] = value;
      ++pop[bin];
      ++index;
    } else {
      ++bin;
    }
  }
  auto population = [&](size_t i) -> size_t { return i <= nbins ? pop[i] : 0; };
  std::vector<double> shells;
  for (size_t i = 0; i <= num_edges; ++i) {
    const size_t prev = i == 0 ? 0 : population(i - 1);
    const size_t f = population(i), next = population(i + 1);
    const auto threshold = static_cast<size_t>((1.0 - peak_isolation) * static_cast<double>(f));
    if (threshold > prev && threshold > next) shells.push_back(mx[i]);
  }
  if (shells.empty()) throw std::invalid_argument("no isolated distance peaks found");
  if (shells.front() != 0.0 && !is_close(shells.front(), 0.0, 1.0e-5, 1.0e-3))
    shells.insert(shells.begin(), 0.0);
  return shells;
}

std::vector<uint32_t> shell_matrix(const std::vector<double> &dm, uint32_t n,
                                   const std::vector<double> &radii, double atol, double rtol) {
  auto find_shell = [&](double x) -> uint32_t {
    if (x < 0) throw std::out_of_range("Invalid distance matrix input");
    if (is_close(x, 0.0, atol, rtol)) return 0;
    for (size_t r = 0; r + 1 < radii.size(); ++r) {
      const double lb = radii[r], ub = radii[r + 1];
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

std::vector<double> prefactors(const std::vector<uint32_t> &sm, uint32_t n,
                               const std::vector<uint32_t> &shells,
                               const std::vector<uint8_t> &species, uint32_t *num_species) {
  if (shells.empty()) throw std::out_of_range("no coordination shells selected");
  // neighbors: shell -> entries over the full N x N matrix, then entries / N stored
  // back into the size_t counter (truncation), structure.h:141-154
  std::map<uint32_t, size_t> neighbors;
  for (uint32_t s : sm) ++neighbors[s];
  for (auto &[shell, count] : neighbors)
    count = static_cast<size_t>(static_cast<double>(count) / static_cast<double>(n));
  std::map<uint8_t, size_t> hist;
  for (uint8_t z : species) ++hist[z];
  std::vector<uint8_t> zs;
  for (auto &[z, c] : hist) zs.push_back(z);
  const uint32_t K = static_cast<uint32_t>(zs.size());
  const size_t S = shells.size();
  std::vector<double> out(S * K * K, 0.0);
  const double N = static_cast<double>(species.size());
  for (size_t s = 0; s < S; ++s) {
    auto it = neighbors.find(shells[s]);
    const size_t count = it == neighbors.end() ? 0 : it->second;
    if (count == 0) continue;  // zeros (the reference logs a warning)
    const double M = static_cast<double>(count);
    for (uint32_t a = 0; a < K; ++a) {
      const double xa = static_cast<double>(hist[zs[a]]) / N;
      for (uint32_t b = a; b < K; ++b) {
        const double xb = static_cast<double>(hist[zs[b]]) / N;
        const double p = 1.0 / (M * xa * xb * N);
        out[(s * K + a) * K + b] = p;
        out[(s * K + b) * K + a] = p;
      }
    }
  }
  if (num_species) *num_species = K;
  return out;
}

void supercell(const Lattice &lat, const double *frac, const uint8_t *species, uint32_t n,
               uint32_t a, uint32_t b, uint32_t c, Lattice *lat_out, double *frac_out,
               uint8_t *species_out) {
  if (a == 0 || b == 0 || c == 0)
    throw std::invalid_argument("There must be at least one copy in each direction");
  // inverse of diag(a, b, c) the way Eigen's 3x3 cofactor inverse evaluates it
  const double sa = a, sb = b, sc = c;
  const double invdet = 1.0 / ((sb * sc) * sa);
  const double inv[3] = {(sb * sc) * invdet, (sc * sa) * invdet, (sa * sb) * invdet};
  size_t site = 0;
  for (uint32_t x = 0; x < a; ++x)
    for (uint32_t y = 0; y < b; ++y)
      for (uint32_t z = 0; z < c; ++z) {
        const double t[3] = {double(x) * inv[0], double(y) * inv[1], double(z) * inv[2]};
        for (uint32_t q = 0; q < n; ++q, ++site) {
          for (int k = 0; k < 3; ++k) frac_out[3 * site + k] = t[k] + frac[3 * q + k] * inv[k];
          species_out[site] = species[q];
        }
      }
  const double s[3] = {sa, sb, sc};
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) lat_out->m[3 * r + k] = lat.m[3 * r + k] * s[k];
}

}  // namespace sqsg
