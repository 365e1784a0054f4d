// Host-side structure geometry: the producer of the scoring kernels' read-only inputs.
// Restates include/sqsgen/core/structure.h (distances, shell radii, shell matrix,
// prefactors, supercell) with the reference's floating-point operation order, so the
// pair lists handed to the device are the ones the reference would build.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <vector>

namespace sqsg {

template <class T> struct LatticeT {
  T m[9];  // row-major, rows are the cell vectors a, b, c
};
using Lattice = LatticeT<double>;

// Every function exists for T = double and T = float (configuration<double> and
// configuration<float>): the reference evaluates the whole geometry in the configuration's
// precision.

// structure.h:34-69: minimum-image distances over the 27 neighbouring cells
template <class T>
std::vector<T> distance_matrix(const LatticeT<T> &lat, const T *frac, uint32_t n);

// structure.h:195-207
template <class T>
std::vector<T> distances_naive(const std::vector<T> &dm, uint32_t n, T atol, T rtol);

// structure.h:209-258 (throws std::invalid_argument below 10 histogram edges)
template <class T>
std::vector<T> distances_histogram(const std::vector<T> &dm, uint32_t n, T bin_width,
                                   T peak_isolation);

// structure.h:71-100
template <class T>
std::vector<uint32_t> shell_matrix(const std::vector<T> &dm, uint32_t n,
                                   const std::vector<T> &radii, T atol, T rtol);

// structure.h:136-192; shells ascending; species are atomic numbers
template <class T>
std::vector<T> prefactors(const std::vector<uint32_t> &sm, uint32_t n,
                          const std::vector<uint32_t> &shells,
                          const std::vector<uint8_t> &species, uint32_t *num_species);

// structure.h:335-366
template <class T>
void supercell(const LatticeT<T> &lat, const T *frac, const uint8_t *species, uint32_t n,
               uint32_t a, uint32_t b, uint32_t c, LatticeT<T> *lat_out, T *frac_out,
               uint8_t *species_out);

constexpr double kEps = 2.220446049250313e-16;  // std::numeric_limits<double>::epsilon()
constexpr float kEpsF = 1.1920929e-7f;            // std::numeric_limits<float>::epsilon()

}  // namespace sqsg
