// Host-side structure geometry: the producer of the scoring kernels' read-only inputs.
// Restates include/sqsgen/core/structure.h (distances, shell radii, shell matrix,
// prefactors, supercell) with the reference's floating-point operation order, so the
// pair lists handed to the device are the ones the reference would build.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <vector>

namespace sqsg {

struct Lattice {
  double m[9];  // row-major, rows are the cell vectors a, b, c
};

// structure.h:34-69: minimum-image distances over the 27 neighbouring cells
std::vector<double> distance_matrix(const Lattice &lat, const double *frac, uint32_t n);

// structure.h:195-207
std::vector<double> distances_naive(const std::vector<double> &dm, uint32_t n, double atol,
                                    double rtol);

// structure.h:209-258 (throws std::invalid_argument below 10 histogram edges)
std::vector<double> distances_histogram(const std::vector<double> &dm, uint32_t n,
                                        double bin_width, double peak_isolation);

// structure.h:71-100
std::vector<uint32_t> shell_matrix(const std::vector<double> &dm, uint32_t n,
                                   const std::vector<double> &radii, double atol, double rtol);

// structure.h:136-192; shells ascending; species are atomic numbers
std::vector<double> prefactors(const std::vector<uint32_t> &sm, uint32_t n,
                               const std::vector<uint32_t> &shells,
                               const std::vector<uint8_t> &species, uint32_t *num_species);

// structure.h:335-366
void supercell(const Lattice &lat, const double *frac, const uint8_t *species, uint32_t n,
               uint32_t a, uint32_t b, uint32_t c, Lattice *lat_out, double *frac_out,
               uint8_t *species_out);

constexpr double kEps = 2.220446049250313e-16;  // std::numeric_limits<double>::epsilon()

}  // namespace sqsg
