/*
 * sqsgen_oracle.c -- CPU restatement of the sqsgenerator scoring/search path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load this library.  Nothing in the
 * product package calls it.
 *
 * Compiled with -ffp-contract=off: the reference is built for plain x86-64
 * without FMA, and result collections key on exact FP objective values.
 * References are file:line in dgehringer/sqsgenerator v0.5.6.
 */
#define _GNU_SOURCE
#include "sqsgen_oracle.h"

#include <math.h>
#include <float.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef unsigned __int128 u128;

/* rapidhash.h:129-131 (first two default secrets are the only ones used) */
static const uint64_t RAPID_SECRET0 = 0x2d358dccaa6c78a5ull;
static const uint64_t RAPID_SECRET1 = 0x8bb84b93962eacc9ull;

/* rapidhash.h:149-197 (RAPIDHASH_FAST, __SIZEOF_INT128__ branch) and 209-212 */
uint64_t or_rapid_mix(uint64_t a, uint64_t b) {
  u128 r = (u128)a * (u128)b;
  return (uint64_t)r ^ (uint64_t)(r >> 64);
}

/* shuffle.h:23-26 */
uint64_t or_rapidrand(uint64_t *state) {
  *state += RAPID_SECRET0;
  return or_rapid_mix(*state, *state ^ RAPID_SECRET1);
}

/* shuffle.h:28-32: low 32 bits of the draw, multiply-shift, no rejection */
uint32_t or_random_bounded(uint32_t range, uint64_t *state) {
  uint32_t x = (uint32_t)or_rapidrand(state);
  uint64_t m = (uint64_t)x * (uint64_t)range;
  return (uint32_t)(m >> 32);
}

/* shuffle.h:39-50: Fisher-Yates per sublattice window, i from w down to 2 */
void or_shuffle_random(uint8_t *conf, const uint32_t *bounds, uint32_t num_bounds,
                       uint64_t *state) {
  for (uint32_t b = 0; b < num_bounds; ++b) {
    uint32_t lo = bounds[2 * b], hi = bounds[2 * b + 1];
    uint32_t w = hi - lo;
    for (uint32_t i = w; i > 1; --i) {
      uint32_t p = or_random_bounded(i, state);
      uint8_t t = conf[lo + i - 1];
      conf[lo + i - 1] = conf[lo + p];
      conf[lo + p] = t;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* permutation.h:17-32: multinomial N!/prod(n_k!) (cpp_int in the reference);
 * built as a product of binomials so every intermediate is exact.
 * Returns 0 when the value fits u128, 1 on overflow. */
int or_num_permutations(const uint8_t *conf, uint32_t n, uint64_t *hi, uint64_t *lo) {
  uint64_t hist[256] = {0};
  for (uint32_t i = 0; i < n; ++i) hist[conf[i]]++;
  u128 total = 1;
  uint64_t placed = 0;
  for (int s = 0; s < 256; ++s) {
    for (uint64_t k = 1; k <= hist[s]; ++k) {
      placed++;
      u128 t;
      if (__builtin_mul_overflow(total, (u128)placed, &t)) return 1;
      total = t / k;
    }
  }
  *hi = (uint64_t)(total >> 64);
  *lo = (uint64_t)total;
  return 0;
}

static uint32_t dense_species_count(const uint8_t *conf, uint32_t n) {
  int seen[256] = {0};
  uint32_t k = 0;
  for (uint32_t i = 0; i < n; ++i)
    if (!seen[conf[i]]) { seen[conf[i]] = 1; k++; }
  return k;
}

/* permutation.h:34-50: 1-based lexicographic rank, built right to left */
void or_rank_permutation(const uint8_t *conf, uint32_t n, uint64_t *hi, uint64_t *lo) {
  uint32_t num_species = dense_species_count(conf, n);
  u128 rank = 1, suffix = 1;
  uint64_t hist[256] = {0};
  for (uint32_t i = 0; i < n; ++i) {
    uint8_t x = conf[(n - 1) - i];
    hist[x]++;
    for (uint32_t j = 0; j < num_species; ++j)
      if (j < x) rank += suffix * hist[j] / hist[x];
    suffix = suffix * (u128)(i + 1) / hist[x];
  }
  *hi = (uint64_t)(rank >> 64);
  *lo = (uint64_t)rank;
}

/* permutation.h:52-78: returns 0, or 2 when rank > total (std::out_of_range) */
int or_unrank_permutation(uint8_t *conf, uint32_t n, uint64_t hi, uint64_t lo) {
  uint64_t counts[256] = {0};
  for (uint32_t i = 0; i < n; ++i) counts[conf[i]]++;
  /* count() is an ordered map: hist holds the counts of the present values in
   * ascending order, and unrank writes hist indices (dense 0..K-1) */
  uint64_t hist[256];
  uint32_t num_species = 0;
  for (int s = 0; s < 256; ++s)
    if (counts[s]) hist[num_species++] = counts[s];
  uint64_t thi, tlo;
  if (or_num_permutations(conf, n, &thi, &tlo)) return 1;
  u128 total = ((u128)thi << 64) | tlo;
  u128 rank = ((u128)hi << 64) | lo;
  if (rank > total) return 2;
  uint32_t k = 0;
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t j = 0; j < num_species; ++j) {
      u128 suffix_count = total * hist[j] / (n - i);
      if (rank <= suffix_count) {
        conf[k] = (uint8_t)j;
        total = suffix_count;
        hist[j]--;
        k++;
        break;
      }
      rank -= suffix_count;
    }
  }
  return 0;
}

/* permutation.h:80-102: lexicographic successor; 0 when already the last */
int or_next_permutation(uint8_t *c, uint32_t n) {
  if (n < 2) return 0;
  long l = (long)n - 1, k = (long)n - 2;
  while (c[k] >= c[k + 1]) {
    k -= 1;
    if (k < 0) return 0;
  }
  while (c[k] >= c[l]) l -= 1;
  uint8_t t = c[k]; c[k] = c[l]; c[l] = t;
  long length = (long)n - (k + 1);
  for (long i = 0; i < length / 2; ++i) {
    long m = k + i + 1, q = (long)n - i - 1;
    t = c[m]; c[m] = c[q]; c[q] = t;
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* numeric.h:11-13 */
static int is_close(double a, double b, double atol, double rtol) {
  return fabs(a - b) <= (atol + rtol * fabs(b));
}

/* structure.h:34-69: minimum image over the 27 neighbouring cells.
 * lattice is row-major with rows a, b, c; cart = frac * lattice. */
void or_distance_matrix(const double *L, const double *frac, uint32_t n, double *out) {
  double *cart = (double *)malloc(sizeof(double) * 3 * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k)
      cart[3 * i + k] = frac[3 * i + 0] * L[0 * 3 + k] + frac[3 * i + 1] * L[1 * 3 + k]
                        + frac[3 * i + 2] * L[2 * 3 + k];
  for (uint64_t q = 0; q < (uint64_t)n * n; ++q) out[q] = DBL_MAX;
  const int axis[3] = {-1, 0, 1};
  for (uint32_t i = 0; i < n; ++i) {
    const double *p1 = cart + 3 * i;
    for (uint32_t j = i; j < n; ++j) {
      const double *p2 = cart + 3 * j;
      for (int iu = 0; iu < 3; ++iu)
        for (int iv = 0; iv < 3; ++iv)
          for (int iw = 0; iw < 3; ++iw) {
            double u = axis[iu], v = axis[iv], w = axis[iw];
            double d[3];
            for (int k = 0; k < 3; ++k) {
              double t = u * L[0 * 3 + k] + v * L[1 * 3 + k] + w * L[2 * 3 + k];
              d[k] = p1[k] - (t + p2[k]);
            }
            double norm = fabs(sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
            if (norm < out[(uint64_t)i * n + j]) {
              out[(uint64_t)i * n + j] = norm;
              if (i != j) out[(uint64_t)j * n + i] = norm;
            }
          }
    }
  }
  free(cart);
}

static int cmp_double(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

/* structure.h:195-207: sorted distinct distances, merged while is_close */
uint32_t or_distances_naive(const double *dm, uint32_t n, double atol, double rtol, double *out,
                            uint32_t cap) {
  uint64_t nn = (uint64_t)n * n;
  double *d = (double *)malloc(sizeof(double) * nn);
  memcpy(d, dm, sizeof(double) * nn);
  qsort(d, nn, sizeof(double), cmp_double);
  /* sorted_vector(range) inserts unique values only (sorted_vector.h:33-36,55-59) */
  uint64_t u = 0;
  for (uint64_t q = 0; q < nn; ++q)
    if (u == 0 || d[q] != d[u - 1]) d[u++] = d[q];
  double *vec = (double *)malloc(sizeof(double) * (u + 1));
  uint32_t m = 1;
  vec[0] = 0.0;
  for (uint64_t q = 0; q < u; ++q) {
    if (is_close(vec[m - 1], d[q], atol, rtol))
      vec[m - 1] = 0.5 * (d[q] + vec[m - 1]);
    else
      vec[m++] = d[q];
  }
  for (uint32_t q = 0; q < m && q < cap; ++q) out[q] = vec[q];
  free(vec);
  free(d);
  return m;
}

/* structure.h:209-258: "peak" radii from an isolated-bin histogram.
 * Returns 0, or 1 when fewer than 10 edges (std::invalid_argument). */
int or_distances_histogram(const double *dm, uint32_t n, double bin_width, double peak_isolation,
                           double *out, uint32_t cap, uint32_t *count) {
  uint64_t nn = (uint64_t)n * n, nd = 0;
  double *d = (double *)malloc(sizeof(double) * (nn ? nn : 1));
  for (uint64_t q = 0; q < nn; ++q)
    if (dm[q] > 0.0 && !is_close(dm[q], 0.0, 1.0e-5, 1.0e-3)) d[nd++] = dm[q];
  if (nd == 0) { free(d); return 1; }
  qsort(d, nd, sizeof(double), cmp_double);
  double min_dist = d[0], max_dist = d[nd - 1];
  uint64_t num_edges = (uint64_t)((max_dist - min_dist) / bin_width) + 2;
  if (num_edges < 10) { free(d); return 1; }
  double *edges = (double *)malloc(sizeof(double) * num_edges);
  for (uint64_t i = 0; i < num_edges; ++i) edges[i] = min_dist + (double)i * bin_width;
  /* per-bin population and maximum; bins are visited in order so the map of
   * the reference holds exactly the bins 0..bin_reached */
  uint64_t nbins = num_edges - 1;
  uint64_t *pop = (uint64_t *)calloc(nbins + 2, sizeof(uint64_t));
  double *mx = (double *)calloc(nbins + 2, sizeof(double));
  uint64_t index = 0, bin = 0;
  while (index < nd && bin < num_edges - 1) {
    double lower = edges[bin], upper = edges[bin + 1], value = d[index];
    if (lower <= value && value < upper) {
      if (pop[bin] == 0 || value > mx[bin]) mx[bin] = value;
      pop[bin]++;
      ++index;
    } else
      ++bin;
  }
  uint32_t m = 0;
  /* for (i = 0; i <= num_edges; i++) with prev = i-1 (absent for i = 0) */
  for (uint64_t i = 0; i <= num_edges; ++i) {
    uint64_t prev = (i == 0) ? 0 : (i - 1 <= nbins ? pop[i - 1] : 0);
    uint64_t f = (i <= nbins) ? pop[i] : 0;
    uint64_t next = (i + 1 <= nbins) ? pop[i + 1] : 0;
    uint64_t threshold = (uint64_t)((1.0 - peak_isolation) * (double)f);
    if (threshold > prev && threshold > next) {
      if (m < cap) out[m] = mx[i];
      m++;
    }
  }
  if (m > 0 && out[0] != 0.0 && !is_close(out[0], 0.0, 1.0e-5, 1.0e-3)) {
    uint32_t top = m < cap ? m : cap - 1;
    for (uint32_t q = top; q > 0; --q) out[q] = out[q - 1];
    out[0] = 0.0;
    m++;
  }
  *count = m;
  free(pop); free(mx); free(edges); free(d);
  return 0;
}

/* structure.h:71-100: shell index of every pair, interval (r_{k-1}, r_k] */
void or_shell_matrix(const double *dm, uint32_t n, const double *radii, uint32_t num_radii,
                     double atol, double rtol, uint32_t *out) {
  for (uint32_t i = 0; i < n; ++i) {
    out[(uint64_t)i * n + i] = 0;
    for (uint32_t j = i + 1; j < n; ++j) {
      double d = dm[(uint64_t)i * n + j];
      uint32_t shell = num_radii;
      if (is_close(d, 0.0, atol, rtol))
        shell = 0;
      else {
        for (uint32_t r = 0; r + 1 < num_radii; ++r) {
          double lb = radii[r], up = radii[r + 1];
          if ((is_close(d, lb, atol, rtol) || d > lb) && (is_close(d, up, atol, rtol) || up > d)) {
            shell = r + 1;
            break;
          }
        }
      }
      out[(uint64_t)i * n + j] = shell;
      out[(uint64_t)j * n + i] = shell;
    }
  }
}

/* structure.h:136-192: pref(s,a,b) = 1/(((M_s x_a) x_b) N).  M_s is the shell's
 * entry count over the full N x N matrix divided by N, stored into a size_t
 * counter (so truncated); shells absent from the matrix give zeros.
 * shells[] must be ascending; returns K (distinct species). */
uint32_t or_compute_prefactors(const uint32_t *sm, uint32_t n, const uint32_t *shells,
                               uint32_t num_shells, const uint8_t *species_z, double *out) {
  uint64_t zcount[256] = {0};
  for (uint32_t i = 0; i < n; ++i) zcount[species_z[i]]++;
  uint8_t zlist[256];
  uint32_t K = 0;
  for (int z = 0; z < 256; ++z)
    if (zcount[z]) zlist[K++] = (uint8_t)z;
  for (uint32_t s = 0; s < num_shells; ++s) {
    uint64_t cnt = 0;
    for (uint64_t q = 0; q < (uint64_t)n * n; ++q) cnt += (sm[q] == shells[s]);
    uint64_t neighbor_count = 0;
    if (cnt > 0) neighbor_count = (uint64_t)((double)cnt / (double)n);
    for (uint32_t a = 0; a < K; ++a)
      for (uint32_t b = a; b < K; ++b) {
        double pref = 0.0;
        if (neighbor_count > 0) {
          double M = (double)neighbor_count;
          double xa = (double)zcount[zlist[a]] / (double)n;
          double xb = (double)zcount[zlist[b]] / (double)n;
          pref = 1.0 / (M * xa * xb * (double)n);
        }
        out[(s * K + a) * K + b] = pref;
        out[(s * K + b) * K + a] = pref;
      }
  }
  return K;
}

/* structure.h:335-366 for a diagonal scale matrix; the inverse follows Eigen's
 * 3x3 cofactor inverse (cofactor * (1/det)). */
void or_supercell(const double *L, const double *frac, const uint8_t *species, uint32_t n,
                  uint32_t a, uint32_t b, uint32_t c, double *Lout, double *fout, uint8_t *sout) {
  double sa = (double)a, sb = (double)b, sc = (double)c;
  double det = (sb * sc) * sa;
  double invdet = 1.0 / det;
  double i0 = (sb * sc) * invdet, i1 = (sc * sa) * invdet, i2 = (sa * sb) * invdet;
  double idiag[3] = {i0, i1, i2};
  uint64_t site = 0;
  for (uint32_t x = 0; x < a; ++x)
    for (uint32_t y = 0; y < b; ++y)
      for (uint32_t z = 0; z < c; ++z) {
        double t[3] = {(double)x * i0, (double)y * i1, (double)z * i2};
        for (uint32_t q = 0; q < n; ++q) {
          for (int k = 0; k < 3; ++k) fout[3 * site + k] = t[k] + frac[3 * q + k] * idiag[k];
          sout[site] = species[q];
          site++;
        }
      }
  double s[3] = {sa, sb, sc};
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) Lout[3 * r + k] = L[3 * r + k] * s[k];
}

/* optimization.h:17-33: scale both triangles of each shell by its weight;
 * weights are given in ascending shell order */
void or_scaled_pair_weights(double *pw, const double *w, uint32_t S, uint32_t K) {
  for (uint32_t s = 0; s < S; ++s)
    for (uint32_t xi = 0; xi < K; ++xi)
      for (uint32_t eta = xi; eta < K; ++eta) {
        pw[(s * K + xi) * K + eta] *= w[s];
        if (xi == eta) continue;
        pw[(s * K + eta) * K + xi] *= w[s];
      }
}

/* structure.h:434-451: upper-triangle pairs whose shell is selected, shell
 * packed to its rank among the selected (ascending) shells */
uint32_t or_pairs(const uint32_t *sm, uint32_t n, const uint32_t *shells, uint32_t S,
                  uint64_t *out, uint32_t cap) {
  uint32_t P = 0;
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j = i + 1; j < n; ++j) {
      uint32_t sh = sm[(uint64_t)i * n + j];
      for (uint32_t s = 0; s < S; ++s)
        if (shells[s] == sh) {
          if (P < cap) {
            out[3 * P + 0] = i;
            out[3 * P + 1] = j;
            out[3 * P + 2] = s;
          }
          P++;
          break;
        }
    }
  return P;
}

/* optimization.h:74-77 */
void or_count_bonds(const uint64_t *pairs, uint32_t P, const uint8_t *species, uint32_t S,
                    uint32_t K, uint64_t *bonds) {
  memset(bonds, 0, sizeof(uint64_t) * S * K * K);
  for (uint32_t p = 0; p < P; ++p) {
    uint64_t i = pairs[3 * p], j = pairs[3 * p + 1], s = pairs[3 * p + 2];
    bonds[(s * K + species[i]) * K + species[j]]++;
  }
}

/* numeric.h:15 */
static double absolute(double a) { return a > 0 ? a : -a; }

/* optimization.h:79-103: FP sum in shell, xi, (diag, eta>xi) order */
double or_compute_objective(const uint64_t *bonds, const double *pref, const double *pw,
                            const double *tgt, uint32_t S, uint32_t K, double *sro) {
  double objective = 0.0;
#define IX(s, a, b) (((s) * K + (a)) * K + (b))
  for (uint32_t s = 0; s < S; ++s)
    for (uint32_t xi = 0; xi < K; ++xi) {
      double sigma = 1.0 - (double)bonds[IX(s, xi, xi)] * pref[IX(s, xi, xi)];
      if (sro) sro[IX(s, xi, xi)] = sigma;
      objective += pw[IX(s, xi, xi)] * absolute(sigma - tgt[IX(s, xi, xi)]);
      for (uint32_t eta = xi + 1; eta < K; ++eta) {
        uint64_t pb = bonds[IX(s, xi, eta)] + bonds[IX(s, eta, xi)];
        double v = 1.0 - (double)pb * pref[IX(s, xi, eta)];
        if (sro) {
          sro[IX(s, xi, eta)] = v;
          sro[IX(s, eta, xi)] = v;
        }
        objective += pw[IX(s, xi, eta)] * absolute(v - tgt[IX(s, xi, eta)]);
      }
    }
#undef IX
  return objective;
}

/* ------------------------------------------------------------------------ */
/* configuration<float> (sqs.h:591-597): the same restatements with T = float.
 * C evaluates float operands in float (FLT_EVAL_METHOD 0 on x86-64 SSE), as the
 * reference's float instantiation does; literals follow the reference's types. */

static int is_close_f(float a, float b, float atol, float rtol) { /* numeric.h:11-13 */
  return fabsf(a - b) <= (atol + rtol * fabsf(b));
}

/* structure.h:34-69 with T = float */
void or_distance_matrix_f32(const float *L, const float *frac, uint32_t n, float *out) {
  float *cart = (float *)malloc(sizeof(float) * 3 * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k)
      cart[3 * i + k] = frac[3 * i + 0] * L[0 * 3 + k] + frac[3 * i + 1] * L[1 * 3 + k]
                        + frac[3 * i + 2] * L[2 * 3 + k];
  for (uint64_t q = 0; q < (uint64_t)n * n; ++q) out[q] = FLT_MAX;
  const int axis[3] = {-1, 0, 1};
  for (uint32_t i = 0; i < n; ++i) {
    const float *p1 = cart + 3 * i;
    for (uint32_t j = i; j < n; ++j) {
      const float *p2 = cart + 3 * j;
      for (int iu = 0; iu < 3; ++iu)
        for (int iv = 0; iv < 3; ++iv)
          for (int iw = 0; iw < 3; ++iw) {
            float u = (float)axis[iu], v = (float)axis[iv], w = (float)axis[iw];
            float d[3];
            for (int k = 0; k < 3; ++k) {
              float t = u * L[0 * 3 + k] + v * L[1 * 3 + k] + w * L[2 * 3 + k];
              d[k] = p1[k] - (t + p2[k]);
            }
            float norm = fabsf(sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
            if (norm < out[(uint64_t)i * n + j]) {
              out[(uint64_t)i * n + j] = norm;
              if (i != j) out[(uint64_t)j * n + i] = norm;
            }
          }
    }
  }
  free(cart);
}

/* structure.h:71-100 with T = float */
void or_shell_matrix_f32(const float *dm, uint32_t n, const float *radii, uint32_t num_radii,
                         float atol, float rtol, uint32_t *out) {
  for (uint32_t i = 0; i < n; ++i) {
    out[(uint64_t)i * n + i] = 0;
    for (uint32_t j = i + 1; j < n; ++j) {
      float d = dm[(uint64_t)i * n + j];
      uint32_t shell = num_radii;
      if (is_close_f(d, 0.0f, atol, rtol))
        shell = 0;
      else {
        for (uint32_t r = 0; r + 1 < num_radii; ++r) {
          float lb = radii[r], up = radii[r + 1];
          if ((is_close_f(d, lb, atol, rtol) || d > lb) &&
              (is_close_f(d, up, atol, rtol) || up > d)) {
            shell = r + 1;
            break;
          }
        }
      }
      out[(uint64_t)i * n + j] = shell;
      out[(uint64_t)j * n + i] = shell;
    }
  }
}

/* structure.h:136-192 with T = float */
uint32_t or_compute_prefactors_f32(const uint32_t *sm, uint32_t n, const uint32_t *shells,
                                   uint32_t num_shells, const uint8_t *species_z, float *out) {
  uint64_t zcount[256] = {0};
  for (uint32_t i = 0; i < n; ++i) zcount[species_z[i]]++;
  uint8_t zlist[256];
  uint32_t K = 0;
  for (int z = 0; z < 256; ++z)
    if (zcount[z]) zlist[K++] = (uint8_t)z;
  for (uint32_t s = 0; s < num_shells; ++s) {
    uint64_t cnt = 0;
    for (uint64_t q = 0; q < (uint64_t)n * n; ++q) cnt += (sm[q] == shells[s]);
    uint64_t neighbor_count = 0;
    if (cnt > 0) neighbor_count = (uint64_t)((float)cnt / (float)n);
    for (uint32_t a = 0; a < K; ++a)
      for (uint32_t b = a; b < K; ++b) {
        float pref = 0.0f;
        if (neighbor_count > 0) {
          float M = (float)neighbor_count;
          float xa = (float)zcount[zlist[a]] / (float)n;
          float xb = (float)zcount[zlist[b]] / (float)n;
          pref = 1.0f / (M * xa * xb * (float)n);
        }
        out[(s * K + a) * K + b] = pref;
        out[(s * K + b) * K + a] = pref;
      }
  }
  return K;
}

/* optimization.h:17-33 with T = float */
void or_scaled_pair_weights_f32(float *pw, const float *w, uint32_t S, uint32_t K) {
  for (uint32_t s = 0; s < S; ++s)
    for (uint32_t xi = 0; xi < K; ++xi)
      for (uint32_t eta = xi; eta < K; ++eta) {
        pw[(s * K + xi) * K + eta] *= w[s];
        if (xi == eta) continue;
        pw[(s * K + eta) * K + xi] *= w[s];
      }
}

/* optimization.h:79-103 with T = float (absolute() of numeric.h:15 in float) */
float or_compute_objective_f32(const uint64_t *bonds, const float *pref, const float *pw,
                               const float *tgt, uint32_t S, uint32_t K, float *sro) {
  float objective = 0.0f;
#define IX(s, a, b) (((s) * K + (a)) * K + (b))
  for (uint32_t s = 0; s < S; ++s)
    for (uint32_t xi = 0; xi < K; ++xi) {
      float sigma = 1.0f - (float)bonds[IX(s, xi, xi)] * pref[IX(s, xi, xi)];
      if (sro) sro[IX(s, xi, xi)] = sigma;
      float a = sigma - tgt[IX(s, xi, xi)];
      objective += pw[IX(s, xi, xi)] * (a > 0 ? a : -a);
      for (uint32_t eta = xi + 1; eta < K; ++eta) {
        uint64_t pb = bonds[IX(s, xi, eta)] + bonds[IX(s, eta, xi)];
        float v = 1.0f - (float)pb * pref[IX(s, xi, eta)];
        if (sro) {
          sro[IX(s, xi, eta)] = v;
          sro[IX(s, eta, xi)] = v;
        }
        float b = v - tgt[IX(s, xi, eta)];
        objective += pw[IX(s, xi, eta)] * (b > 0 ? b : -b);
      }
    }
#undef IX
  return objective;
}

/* ------------------------------------------------------------------------ */
/* Driver restatement: sqs.h:299-467 worker loop, results.h:42-99 collection. */

typedef struct {
  double objective;
  uint8_t *species;      /* concatenated per sublattice */
  double *sub_objectives;
} or_entry;

static struct {
  pthread_mutex_t lock;
  or_entry *entries;
  uint64_t num_entries, cap_entries;
  double *levels;        /* sorted distinct objectives (results.h:97) */
  uint64_t *level_count; /* set size per level (results.h:63-69) */
  uint64_t num_levels, cap_levels;
  uint64_t *table;       /* open-addressing set of entry index + 1, keyed on (objective, species) */
  uint64_t table_cap;
  uint32_t blob;         /* species bytes per entry */
  uint32_t nsub;
} g_coll = {PTHREAD_MUTEX_INITIALIZER, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

static uint64_t entry_hash(double obj, const uint8_t *species, uint32_t blob) {
  uint64_t h = 0xcbf29ce484222325ull, b;
  memcpy(&b, &obj, 8);
  h = (h ^ b) * 0x100000001b3ull;
  for (uint32_t i = 0; i < blob; ++i) h = (h ^ species[i]) * 0x100000001b3ull;
  return h ^ (h >> 29);
}

static void coll_reset(uint32_t blob, uint32_t nsub) {
  for (uint64_t e = 0; e < g_coll.num_entries; ++e) {
    free(g_coll.entries[e].species);
    free(g_coll.entries[e].sub_objectives);
  }
  free(g_coll.entries);
  free(g_coll.levels);
  free(g_coll.level_count);
  free(g_coll.table);
  g_coll.entries = 0; g_coll.levels = 0; g_coll.level_count = 0;
  g_coll.num_entries = g_coll.cap_entries = g_coll.num_levels = g_coll.cap_levels = 0;
  g_coll.table_cap = 1024;
  g_coll.table = (uint64_t *)calloc(g_coll.table_cap, sizeof(uint64_t));
  g_coll.blob = blob;
  g_coll.nsub = nsub;
}

static uint64_t level_pos(double obj) { /* first level >= obj */
  uint64_t lo = 0, hi = g_coll.num_levels;
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    if (g_coll.levels[mid] < obj) lo = mid + 1; else hi = mid;
  }
  return lo;
}

static uint64_t coll_count_for(double obj) { /* results.h:63-69 */
  uint64_t pos = level_pos(obj);
  return (pos < g_coll.num_levels && g_coll.levels[pos] == obj) ? g_coll.level_count[pos] : 0;
}

static double coll_nth_best(uint64_t n) { /* results.h:78-82 */
  return n >= g_coll.num_levels ? INFINITY : g_coll.levels[n];
}

static void table_put(uint64_t idx) {
  const or_entry *en = &g_coll.entries[idx];
  uint64_t mask = g_coll.table_cap - 1;
  uint64_t h = entry_hash(en->objective, en->species, g_coll.blob) & mask;
  while (g_coll.table[h]) h = (h + 1) & mask;
  g_coll.table[h] = idx + 1;
}

static void coll_insert(double obj, const uint8_t *species, const double *subobj) {
  /* results.h:46-53: a set per objective keyed on the species */
  uint64_t mask = g_coll.table_cap - 1;
  uint64_t h = entry_hash(obj, species, g_coll.blob) & mask;
  while (g_coll.table[h]) {
    const or_entry *en = &g_coll.entries[g_coll.table[h] - 1];
    if (en->objective == obj && memcmp(en->species, species, g_coll.blob) == 0) return;
    h = (h + 1) & mask;
  }
  if (g_coll.num_entries == g_coll.cap_entries) {
    g_coll.cap_entries = g_coll.cap_entries ? 2 * g_coll.cap_entries : 64;
    g_coll.entries = (or_entry *)realloc(g_coll.entries, sizeof(or_entry) * g_coll.cap_entries);
  }
  uint64_t idx = g_coll.num_entries++;
  or_entry *en = &g_coll.entries[idx];
  en->objective = obj;
  en->species = (uint8_t *)malloc(g_coll.blob);
  memcpy(en->species, species, g_coll.blob);
  en->sub_objectives = (double *)malloc(sizeof(double) * g_coll.nsub);
  memcpy(en->sub_objectives, subobj, sizeof(double) * g_coll.nsub);
  g_coll.table[h] = idx + 1;
  if (2 * g_coll.num_entries > g_coll.table_cap) {
    free(g_coll.table);
    g_coll.table_cap *= 2;
    g_coll.table = (uint64_t *)calloc(g_coll.table_cap, sizeof(uint64_t));
    for (uint64_t e = 0; e < g_coll.num_entries; ++e) table_put(e);
  }
  uint64_t pos = level_pos(obj);
  if (pos < g_coll.num_levels && g_coll.levels[pos] == obj) {
    g_coll.level_count[pos]++;
    return;
  }
  if (g_coll.num_levels == g_coll.cap_levels) {
    g_coll.cap_levels = g_coll.cap_levels ? 2 * g_coll.cap_levels : 64;
    g_coll.levels = (double *)realloc(g_coll.levels, sizeof(double) * g_coll.cap_levels);
    g_coll.level_count = (uint64_t *)realloc(g_coll.level_count, sizeof(uint64_t) * g_coll.cap_levels);
  }
  memmove(g_coll.levels + pos + 1, g_coll.levels + pos,
          sizeof(double) * (g_coll.num_levels - pos));
  memmove(g_coll.level_count + pos + 1, g_coll.level_count + pos,
          sizeof(uint64_t) * (g_coll.num_levels - pos));
  g_coll.levels[pos] = obj;
  g_coll.level_count[pos] = 1;
  g_coll.num_levels++;
}

typedef struct {
  const or_run_config *cfg;
  uint64_t num_blocks, block_size, remainder;
  _Atomic uint64_t next_block;
  _Atomic uint64_t search_bits;  /* std::atomic<T> _search_objective */
  _Atomic uint64_t best_bits;
  _Atomic uint64_t finished;
  _Atomic uint64_t *rng;         /* shared shuffler state, one per sublattice */
  uint32_t blob;
} run_ctx;

static double bits2d(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t d2bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

/* shared-state draw: a relaxed load/store pair, racy like the reference's
 * shared shuffler (sqs.h:320,348) but without undefined behaviour */
static uint32_t bounded_shared(uint32_t range, _Atomic uint64_t *st) {
  uint64_t s = atomic_load_explicit(st, memory_order_relaxed);
  uint32_t r = or_random_bounded(range, &s);
  atomic_store_explicit(st, s, memory_order_relaxed);
  return r;
}

static void shuffle_window(uint8_t *conf, const or_sublattice *sl, _Atomic uint64_t *shared,
                           uint64_t *local) {
  for (uint32_t b = 0; b < sl->num_bounds; ++b) {
    uint32_t lo = sl->bounds[2 * b], hi = sl->bounds[2 * b + 1];
    for (uint32_t i = hi - lo; i > 1; --i) {
      uint32_t p = shared ? bounded_shared(i, shared) : or_random_bounded(i, local);
      uint8_t t = conf[lo + i - 1];
      conf[lo + i - 1] = conf[lo + p];
      conf[lo + p] = t;
    }
  }
}

/* shuffle.h:69-88: unrank inside the (single) window, remapping species to
 * local dense indices and back */
static int unrank_window(uint8_t *conf, const or_sublattice *sl, u128 rank) {
  uint32_t lo = sl->bounds[0], hi = sl->bounds[1];
  uint32_t w = hi - lo;
  if (lo == 0 && hi == sl->num_sites)
    return or_unrank_permutation(conf, sl->num_sites, (uint64_t)(rank >> 64), (uint64_t)rank);
  int present[256] = {0};
  for (uint32_t i = lo; i < hi; ++i) present[conf[i]] = 1;
  uint8_t fwd[256], rev[256];
  uint32_t k = 0;
  for (int s = 0; s < 256; ++s)
    if (present[s]) { fwd[s] = (uint8_t)k; rev[k] = (uint8_t)s; k++; }
  uint8_t *tmp = (uint8_t *)malloc(w ? w : 1);
  for (uint32_t i = 0; i < w; ++i) tmp[i] = fwd[conf[lo + i]];
  int rc = or_unrank_permutation(tmp, w, (uint64_t)(rank >> 64), (uint64_t)rank);
  for (uint32_t i = 0; i < w; ++i) conf[lo + i] = rev[tmp[i]];
  free(tmp);
  return rc;
}

static void *worker(void *arg) {
  run_ctx *ctx = (run_ctx *)arg;
  const or_run_config *cfg = ctx->cfg;
  uint32_t nsub = cfg->num_sublattices;
  uint64_t *bonds[16];
  double *sro[16];
  uint8_t *species = (uint8_t *)malloc(ctx->blob);
  uint32_t off[17];
  off[0] = 0;
  for (uint32_t q = 0; q < nsub; ++q) {
    const or_sublattice *sl = &cfg->sublattices[q];
    off[q + 1] = off[q] + sl->num_sites;
    bonds[q] = (uint64_t *)malloc(sizeof(uint64_t) * sl->num_shells * sl->num_species * sl->num_species);
    sro[q] = (double *)malloc(sizeof(double) * sl->num_shells * sl->num_species * sl->num_species);
  }
  uint64_t local_rng[16];
  double subobj[16];
  for (;;) {
    uint64_t blk = atomic_fetch_add(&ctx->next_block, 1);
    if (blk >= ctx->num_blocks) break;
    /* BS::thread_pool blocks: equal sizes, the first `remainder` one larger */
    uint64_t rstart = blk * ctx->block_size + (blk < ctx->remainder ? blk : ctx->remainder);
    uint64_t rend = rstart + ctx->block_size + (blk < ctx->remainder ? 1 : 0);
    for (uint32_t q = 0; q < nsub; ++q) {
      memcpy(species + off[q], cfg->sublattices[q].species_packed, cfg->sublattices[q].num_sites);
      /* private streams for the per-thread variant: jump ahead by the block start */
      local_rng[q] = cfg->sublattices[q].seed + rstart * 0x9e3779b97f4a7c15ull;
    }
    if (cfg->systematic) unrank_window(species, &cfg->sublattices[0], (u128)rstart + 1);
    int stalled = 0; /* interact: a max_results skip left the configuration unchanged */
    for (uint64_t it = rstart; it < rend; ++it) {
      double objective_value;
      if (cfg->counter_stream && !cfg->systematic && !stalled) {
        for (uint32_t q = 0; q < nsub; ++q) {
          const or_sublattice *sl = &cfg->sublattices[q];
          memcpy(species + off[q], sl->species_packed, sl->num_sites);
          if (!cfg->split && it == rstart) continue; /* sqs.h:374,400-402 */
          uint64_t draws = 0;
          for (uint32_t b = 0; b < sl->num_bounds; ++b)
            draws += sl->bounds[2 * b + 1] - sl->bounds[2 * b] - 1;
          uint64_t st = sl->seed + it * draws * RAPID_SECRET0;
          shuffle_window(species + off[q], sl, 0, &st);
        }
      }
      if (!cfg->split) {
        const or_sublattice *sl = &cfg->sublattices[0];
        or_count_bonds(sl->pairs, sl->num_pairs, species, sl->num_shells, sl->num_species, bonds[0]);
        objective_value = or_compute_objective(bonds[0], sl->prefactors, sl->pair_weights,
                                               sl->target, sl->num_shells, sl->num_species, sro[0]);
        subobj[0] = objective_value;
      } else {
        for (uint32_t q = 0; q < nsub; ++q) {
          const or_sublattice *sl = &cfg->sublattices[q];
          if (!cfg->counter_stream)
            shuffle_window(species + off[q], sl, cfg->shared_rng ? &ctx->rng[q] : 0,
                           &local_rng[q]);
          or_count_bonds(sl->pairs, sl->num_pairs, species + off[q], sl->num_shells,
                         sl->num_species, bonds[q]);
          subobj[q] = or_compute_objective(bonds[q], sl->prefactors, sl->pair_weights, sl->target,
                                           sl->num_shells, sl->num_species, sro[q]);
        }
        objective_value = 0.0; /* helpers::sum: fold_left from T(0) (fold.h:51-53) */
        for (uint32_t q = 0; q < nsub; ++q) objective_value += subobj[q];
      }
      double search = bits2d(atomic_load_explicit(&ctx->search_bits, memory_order_relaxed));
      if (objective_value <= search) {
        pthread_mutex_lock(&g_coll.lock);
        int skip = cfg->max_results_per_objective >= 0
                   && coll_count_for(objective_value) > (uint64_t)cfg->max_results_per_objective;
        if (!skip) {
          coll_insert(objective_value, species, subobj);
          double nb = coll_nth_best(cfg->keep);
          /* sqs.h:203-206 */
          if (objective_value < bits2d(atomic_load(&ctx->best_bits)))
            atomic_store(&ctx->best_bits, d2bits(objective_value));
          if (nb < bits2d(atomic_load(&ctx->search_bits))) atomic_store(&ctx->search_bits, d2bits(nb));
        }
        pthread_mutex_unlock(&g_coll.lock);
        if (skip) { /* sqs.h:423-426: skips the shuffle below as well */
          if (!cfg->split) stalled = 1;
          continue;
        }
      }
      if (!cfg->split) {
        const or_sublattice *sl = &cfg->sublattices[0];
        if (cfg->systematic) {
          uint32_t lo = sl->bounds[0], hi = sl->bounds[1];
          or_next_permutation(species + lo, hi - lo);
        } else if (!cfg->counter_stream)
          shuffle_window(species, sl, cfg->shared_rng ? &ctx->rng[0] : 0, &local_rng[0]);
      }
    }
    atomic_fetch_add(&ctx->finished, rend - rstart);
  }
  for (uint32_t q = 0; q < nsub; ++q) { free(bonds[q]); free(sro[q]); }
  free(species);
  return 0;
}

int or_run(const or_run_config *cfg, or_run_stats *stats) {
  if (cfg->num_sublattices == 0 || cfg->num_sublattices > 16) return 1;
  if (cfg->chunk_size == 0 || cfg->iterations == 0) return 1;
  uint32_t blob = 0;
  for (uint32_t q = 0; q < cfg->num_sublattices; ++q) blob += cfg->sublattices[q].num_sites;
  coll_reset(blob, cfg->num_sublattices);
  run_ctx ctx;
  memset(&ctx, 0, sizeof(ctx));
  ctx.cfg = cfg;
  ctx.blob = blob;
  uint64_t total = cfg->iterations;
  uint64_t nb = total / cfg->chunk_size;
  if (nb == 0) nb = 1;
  if (nb > total) nb = total;
  ctx.num_blocks = nb;
  ctx.block_size = total / nb;
  ctx.remainder = total % nb;
  atomic_store(&ctx.search_bits, d2bits(DBL_MAX));  /* sqs.h:259-260 */
  atomic_store(&ctx.best_bits, d2bits(DBL_MAX));
  _Atomic uint64_t rng[16];
  for (uint32_t q = 0; q < cfg->num_sublattices; ++q) atomic_store(&rng[q], cfg->sublattices[q].seed);
  ctx.rng = rng;
  uint32_t T = cfg->num_threads ? cfg->num_threads : 1;
  pthread_t th[256];
  if (T > 256) T = 256;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (uint32_t t = 0; t < T; ++t) pthread_create(&th[t], 0, worker, &ctx);
  for (uint32_t t = 0; t < T; ++t) pthread_join(th[t], 0);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (stats) {
    stats->finished = atomic_load(&ctx.finished);
    stats->best_objective = bits2d(atomic_load(&ctx.best_bits));
    stats->num_results = g_coll.num_entries;
    stats->num_objectives = g_coll.num_levels;
    stats->seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  }
  return 0;
}

uint64_t or_result_count(void) { return g_coll.num_entries; }

static int cmp_entry(const void *a, const void *b) {
  const or_entry *x = (const or_entry *)a, *y = (const or_entry *)b;
  if (x->objective < y->objective) return -1;
  if (x->objective > y->objective) return 1;
  return memcmp(x->species, y->species, g_coll.blob);
}

int or_result_get(uint64_t r, double *objective, uint8_t *species, double *sub_objectives) {
  static uint64_t sorted_for = (uint64_t)-1;
  if (r >= g_coll.num_entries) return 1;
  if (sorted_for != g_coll.num_entries || r == 0) {
    qsort(g_coll.entries, g_coll.num_entries, sizeof(or_entry), cmp_entry);
    sorted_for = g_coll.num_entries;
  }
  *objective = g_coll.entries[r].objective;
  if (species) memcpy(species, g_coll.entries[r].species, g_coll.blob);
  if (sub_objectives)
    memcpy(sub_objectives, g_coll.entries[r].sub_objectives, sizeof(double) * g_coll.nsub);
  return 0;
}
