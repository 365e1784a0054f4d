/*
 * sqsgen_oracle.h -- CPU restatement of the sqsgenerator scoring/search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be linked,
 * loaded or executed only by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py.  The product (sqsgenerator_amd) never calls it.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the dgehringer/sqsgenerator source tree, v0.5.6).  Layouts: cubes are
 * row-major [S][K][K] (the reference stores Eigen ColMajor tensors; layout does
 * not change any arithmetic), pair records are 24 bytes {i, j, shell} like the
 * reference's atom_pair<usize_t> (core/structure.h:21-27), bond counters are
 * 64-bit like cube_t<usize_t>.
 *
 * Pinning status (tests/test_oracle.py unless noted):
 *   or_rapid_mix           golden vectors of the reference's own rapid_mix, compiled from
 *                          rapidhash.h into oracle/_ref (tests/golden/rapid_mix.json)
 *   or_rapidrand           golden draw streams from the same build; counter property
 *   or_random_bounded      formula and range; Fisher-Yates uniformity (chi-square) and
 *   or_shuffle_random      window independence.  The exact trajectory of a reference MC
 *                          run (threads sharing the RNG state, block split of the thread
 *                          pool, std::sort site order) is PARITY UNPINNED: MC results are
 *                          checked by re-scoring instead.
 *   or_num/rank/unrank/next_permutation  known answers (C(32,16), first/last ranks),
 *                          brute-force lexicographic positions, unrank/rank round trip
 *   or_distance_matrix, or_shell_matrix, or_supercell  the reference's fixture case1
 *                          (its tests/assets data, in tests/golden/case1.*.json); FCC
 *                          coordination numbers
 *   or_distances_naive     FCC 3x3x3 naive distances and coordination numbers
 *   or_distances_histogram equals the naive radii on perfect FCC 2x2x2 / 3x3x3 lattices
 *                          (no golden vector of the reference's own histogram exists)
 *   or_compute_prefactors, or_scaled_pair_weights  closed forms
 *   or_count_bonds         brute-force dense contraction X^T U X; per-shell totals = pairs
 *   or_compute_objective   closed forms, zero weights, unit prefactors, summation order
 *   or_*_f32               configuration<float>: the same statements with T = float; pinned
 *                          by the same closed forms and invariants in float, by agreement
 *                          with the double functions to float rounding, and by the FCC
 *                          coordination numbers from the float shell matrix
 *   or_run                 systematic: all 12870 results with zero pair weights; keep
 *                          levels hold every tie; max_results_per_objective allows max+1;
 *                          the whole collection (every level, never evicted, the stall of
 *                          a block after a max_results skip) equals a plain-Python
 *                          statement of the one-thread worker loop, in systematic and in
 *                          counter_stream mode, over BS::thread_pool's block split.
 *                          counter_stream: with keep above the candidate count the pack is
 *                          the candidate set built from the golden-pinned shuffle and draw
 *                          stream.  The reference's own MC trajectory (cumulative shuffles
 *                          of a shared RNG state) stays PARITY UNPINNED.
 */
#ifndef SQSGEN_ORACLE_H
#define SQSGEN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: core/helpers/rapidhash.h:129-131,149-212, core/shuffle.h:23-50 ---- */
uint64_t or_rapid_mix(uint64_t a, uint64_t b);
uint64_t or_rapidrand(uint64_t *state);
uint32_t or_random_bounded(uint32_t range, uint64_t *state);
void or_shuffle_random(uint8_t *conf, const uint32_t *bounds, uint32_t num_bounds,
                       uint64_t *state);

/* ---- multiset permutations: core/permutation.h:17-102 (rank_t as u128 hi:lo) ---- */
int or_num_permutations(const uint8_t *conf, uint32_t n, uint64_t *hi, uint64_t *lo);
void or_rank_permutation(const uint8_t *conf, uint32_t n, uint64_t *hi, uint64_t *lo);
int or_unrank_permutation(uint8_t *conf, uint32_t n, uint64_t hi, uint64_t lo);
int or_next_permutation(uint8_t *conf, uint32_t n);

/* ---- geometry: core/structure.h:34-258,335-366 ---- */
void or_distance_matrix(const double *lattice, const double *frac, uint32_t n, double *out);
uint32_t or_distances_naive(const double *dm, uint32_t n, double atol, double rtol, double *out,
                            uint32_t cap);
int or_distances_histogram(const double *dm, uint32_t n, double bin_width, double peak_isolation,
                           double *out, uint32_t cap, uint32_t *count);
void or_shell_matrix(const double *dm, uint32_t n, const double *radii, uint32_t num_radii,
                     double atol, double rtol, uint32_t *out);
uint32_t or_compute_prefactors(const uint32_t *sm, uint32_t n, const uint32_t *shells,
                               uint32_t num_shells, const uint8_t *species_z, double *out);
void or_supercell(const double *lattice, const double *frac, const uint8_t *species, uint32_t n,
                  uint32_t a, uint32_t b, uint32_t c, double *lattice_out, double *frac_out,
                  uint8_t *species_out);

/* ---- optimisation inputs: core/optimization.h:17-72, core/structure.h:434-451 ---- */
void or_scaled_pair_weights(double *pw, const double *shell_weights_sorted, uint32_t num_shells,
                            uint32_t num_species);
uint32_t or_pairs(const uint32_t *sm, uint32_t n, const uint32_t *shells, uint32_t num_shells,
                  uint64_t *out, uint32_t cap);

/* ---- hot path: core/optimization.h:74-103 ---- */
void or_count_bonds(const uint64_t *pairs, uint32_t num_pairs, const uint8_t *species,
                    uint32_t num_shells, uint32_t num_species, uint64_t *bonds);
double or_compute_objective(const uint64_t *bonds, const double *prefactors,
                            const double *pair_weights, const double *target, uint32_t num_shells,
                            uint32_t num_species, double *sro);

/* ---- configuration<float> (T = float) restatements of the same functions ---- */
void or_distance_matrix_f32(const float *lattice, const float *frac, uint32_t n, float *out);
void or_shell_matrix_f32(const float *dm, uint32_t n, const float *radii, uint32_t num_radii,
                         float atol, float rtol, uint32_t *out);
uint32_t or_compute_prefactors_f32(const uint32_t *sm, uint32_t n, const uint32_t *shells,
                                   uint32_t num_shells, const uint8_t *species_z, float *out);
void or_scaled_pair_weights_f32(float *pw, const float *shell_weights_sorted,
                                uint32_t num_shells, uint32_t num_species);
float or_compute_objective_f32(const uint64_t *bonds, const float *prefactors,
                               const float *pair_weights, const float *target,
                               uint32_t num_shells, uint32_t num_species, float *sro);

/* ---- driver: sqs.h:226-238,299-467, core/results.h:42-99 ---- */
typedef struct {
  uint32_t num_sites;
  uint32_t num_species;
  uint32_t num_shells;
  uint32_t num_pairs;
  const uint64_t *pairs;          /* num_pairs x {i, j, shell} */
  const uint8_t *species_packed;  /* num_sites */
  const uint32_t *bounds;         /* num_bounds x {lo, hi} */
  uint32_t num_bounds;
  const double *prefactors;       /* [S][K][K] */
  const double *pair_weights;     /* [S][K][K], already scaled by the shell weights */
  const double *target;           /* [S][K][K] */
  uint64_t seed;
} or_sublattice;

typedef struct {
  int systematic;        /* 0: random, 1: systematic (interact, one sublattice) */
  int split;             /* 0: interact, 1: split */
  uint32_t num_sublattices;
  const or_sublattice *sublattices;
  uint64_t iterations;
  uint64_t chunk_size;
  uint64_t keep;
  int64_t max_results_per_objective; /* -1: unlimited */
  uint32_t num_threads;
  int shared_rng;        /* 1: one RNG state shared by all threads (reference behaviour) */
  int collect;           /* 1: keep the result collection (tests); 0: count only (timing) */
  /* 1: Monte Carlo candidate `it` is the sorted configuration shuffled with the draws
   * [it * D, (it + 1) * D) of the seed's stream (D = draws per candidate), the first of an
   * interact block left unshuffled -- the counter-based candidate definition of the device
   * path; 0: the reference's cumulative shuffles of one running configuration */
  int counter_stream;
} or_run_config;

typedef struct {
  uint64_t finished;
  double best_objective;
  uint64_t num_results;     /* entries in the collection after remove_duplicates */
  uint64_t num_objectives;  /* distinct objective levels in the collection */
  double seconds;           /* wall time of the candidate loop */
} or_run_stats;

/* Runs the optimisation loop; results stay in an internal collection that
 * or_result_* reads until the next or_run call (not thread-safe). */
int or_run(const or_run_config *cfg, or_run_stats *stats);
uint64_t or_result_count(void);
/* result r (sorted by objective, then by insertion): objective, per-sublattice
 * species concatenated (sum of num_sites bytes), per-sublattice objective */
int or_result_get(uint64_t r, double *objective, uint8_t *species, double *sub_objectives);

#ifdef __cplusplus
}
#endif
#endif
