// extern "C" boundary: status codes instead of exceptions, thread-local error text.
#include <hip/hip_runtime_api.h>

#include <cstring>
#include <new>
#include <string>

#include "../../include/sqsgen_amd.h"
#include "geometry.hpp"
#include "job.hpp"
#include "kernels.hpp"
#include "permutation.hpp"

namespace {

thread_local std::string g_last_error;

template <class Fn> int guard(Fn &&fn) {
  try {
    fn();
    return SQSG_OK;
  } catch (const sqsg::CancelledError &e) {
    g_last_error = e.what();
    return SQSG_ERR_CANCELLED;
  } catch (const sqsg::DeviceError &e) {
    g_last_error = e.what();
    return SQSG_ERR_DEVICE;
  } catch (const std::bad_alloc &) {
    g_last_error = "out of memory";
    return SQSG_ERR_NO_MEMORY;
  } catch (const std::out_of_range &e) {
    g_last_error = e.what();
    return SQSG_ERR_OUT_OF_RANGE;
  } catch (const std::invalid_argument &e) {
    g_last_error = e.what();
    return SQSG_ERR_INVALID_ARGUMENT;
  } catch (const std::domain_error &e) {
    g_last_error = e.what();
    return SQSG_ERR_DOMAIN;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return SQSG_ERR_RUNTIME;
  } catch (...) {
    g_last_error = "unknown error";
    return SQSG_ERR_RUNTIME;
  }
}

void need(const void *p, const char *name) {
  if (!p) throw std::invalid_argument(std::string(name) + " must not be NULL");
}

void copy_str(const std::string &s, char *out, size_t cap) {
  if (!out) return;
  if (s.size() + 1 > cap) throw std::out_of_range("output buffer too small");
  std::memcpy(out, s.c_str(), s.size() + 1);
}

}  // namespace

namespace sqsg {
int guarded(const std::function<void()> &fn) { return guard(fn); }
}  // namespace sqsg

extern "C" {

const char *sqsg_last_error(void) { return g_last_error.c_str(); }

int sqsg_version(void) { return 10000; }

int sqsg_device_info(char *name, size_t cap, int *compute_units) {
  return guard([&] {
    int dev = 0;
    SQSG_HIP(hipSetDevice(0));
    SQSG_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SQSG_HIP(hipGetDeviceProperties(&prop, dev));
    if (name) copy_str(std::string(prop.name) + " " + prop.gcnArchName, name, cap);
    if (compute_units) *compute_units = prop.multiProcessorCount;
  });
}

int sqsg_device_synchronize(void) { return guard([] { SQSG_HIP(hipDeviceSynchronize()); }); }

}  // extern "C"

namespace {
// the geometry entry points for T = double and T = float (structure<T>)
template <class T> int distance_matrix_t(const T *lat, const T *frac, uint32_t n, T *out) {
  return guard([&] {
    need(lat, "lattice"), need(frac, "frac"), need(out, "out");
    sqsg::LatticeT<T> L;
    std::memcpy(L.m, lat, sizeof(L.m));
    const auto d = sqsg::distance_matrix<T>(L, frac, n);
    std::memcpy(out, d.data(), d.size() * sizeof(T));
  });
}

template <class T>
int distances_naive_t(const T *dm, uint32_t n, T atol, T rtol, T *out, uint32_t cap,
                      uint32_t *count) {
  return guard([&] {
    need(dm, "dm"), need(count, "count");
    const auto d = sqsg::distances_naive<T>(std::vector<T>(dm, dm + size_t(n) * n), n, atol, rtol);
    *count = static_cast<uint32_t>(d.size());
    if (out) std::memcpy(out, d.data(), std::min<size_t>(cap, d.size()) * sizeof(T));
  });
}

template <class T>
int distances_histogram_t(const T *dm, uint32_t n, T bin_width, T peak_isolation, T *out,
                          uint32_t cap, uint32_t *count) {
  return guard([&] {
    need(dm, "dm"), need(count, "count");
    const auto d = sqsg::distances_histogram<T>(std::vector<T>(dm, dm + size_t(n) * n), n,
                                                bin_width, peak_isolation);
    *count = static_cast<uint32_t>(d.size());
    if (out) std::memcpy(out, d.data(), std::min<size_t>(cap, d.size()) * sizeof(T));
  });
}

template <class T>
int shell_matrix_t(const T *dm, uint32_t n, const T *radii, uint32_t num_radii, T atol, T rtol,
                   uint32_t *out) {
  return guard([&] {
    need(dm, "dm"), need(radii, "radii"), need(out, "out");
    const auto sm = sqsg::shell_matrix<T>(std::vector<T>(dm, dm + size_t(n) * n), n,
                                          std::vector<T>(radii, radii + num_radii), atol, rtol);
    std::memcpy(out, sm.data(), sm.size() * sizeof(uint32_t));
  });
}

template <class T>
int prefactors_t(const uint32_t *sm, uint32_t n, const uint32_t *shells, uint32_t num_shells,
                 const uint8_t *species_z, T *out) {
  return guard([&] {
    need(sm, "shell_matrix"), need(shells, "shells"), need(species_z, "species"), need(out, "out");
    uint32_t K = 0;
    const auto p = sqsg::prefactors<T>(std::vector<uint32_t>(sm, sm + size_t(n) * n), n,
                                       std::vector<uint32_t>(shells, shells + num_shells),
                                       std::vector<uint8_t>(species_z, species_z + n), &K);
    std::memcpy(out, p.data(), p.size() * sizeof(T));
  });
}

template <class T>
int supercell_t(const T *lat, const T *frac, const uint8_t *species, uint32_t n, uint32_t a,
                uint32_t b, uint32_t c, T *lattice_out, T *frac_out, uint8_t *species_out) {
  return guard([&] {
    need(lat, "lattice"), need(frac, "frac"), need(species, "species");
    need(lattice_out, "lattice_out"), need(frac_out, "frac_out"), need(species_out, "species_out");
    sqsg::LatticeT<T> L, lo;
    std::memcpy(L.m, lat, sizeof(L.m));
    sqsg::supercell<T>(L, frac, species, n, a, b, c, &lo, frac_out, species_out);
    std::memcpy(lattice_out, lo.m, sizeof(lo.m));
  });
}
}  // namespace

extern "C" {

int sqsg_distance_matrix(const double *lat, const double *frac, uint32_t n, double *out) {
  return distance_matrix_t<double>(lat, frac, n, out);
}
int sqsg_distances_naive(const double *dm, uint32_t n, double atol, double rtol, double *out,
                         uint32_t cap, uint32_t *count) {
  return distances_naive_t<double>(dm, n, atol, rtol, out, cap, count);
}
int sqsg_distances_histogram(const double *dm, uint32_t n, double bin_width,
                             double peak_isolation, double *out, uint32_t cap, uint32_t *count) {
  return distances_histogram_t<double>(dm, n, bin_width, peak_isolation, out, cap, count);
}
int sqsg_shell_matrix(const double *dm, uint32_t n, const double *radii, uint32_t num_radii,
                      double atol, double rtol, uint32_t *out) {
  return shell_matrix_t<double>(dm, n, radii, num_radii, atol, rtol, out);
}
int sqsg_prefactors(const uint32_t *sm, uint32_t n, const uint32_t *shells, uint32_t num_shells,
                    const uint8_t *species_z, double *out) {
  return prefactors_t<double>(sm, n, shells, num_shells, species_z, out);
}
int sqsg_supercell(const double *lat, const double *frac, const uint8_t *species, uint32_t n,
                   uint32_t a, uint32_t b, uint32_t c, double *lattice_out, double *frac_out,
                   uint8_t *species_out) {
  return supercell_t<double>(lat, frac, species, n, a, b, c, lattice_out, frac_out, species_out);
}

int sqsg_distance_matrix_f32(const float *lat, const float *frac, uint32_t n, float *out) {
  return distance_matrix_t<float>(lat, frac, n, out);
}
int sqsg_distances_naive_f32(const float *dm, uint32_t n, float atol, float rtol, float *out,
                             uint32_t cap, uint32_t *count) {
  return distances_naive_t<float>(dm, n, atol, rtol, out, cap, count);
}
int sqsg_distances_histogram_f32(const float *dm, uint32_t n, float bin_width,
                                 float peak_isolation, float *out, uint32_t cap,
                                 uint32_t *count) {
  return distances_histogram_t<float>(dm, n, bin_width, peak_isolation, out, cap, count);
}
int sqsg_shell_matrix_f32(const float *dm, uint32_t n, const float *radii, uint32_t num_radii,
                          float atol, float rtol, uint32_t *out) {
  return shell_matrix_t<float>(dm, n, radii, num_radii, atol, rtol, out);
}
int sqsg_prefactors_f32(const uint32_t *sm, uint32_t n, const uint32_t *shells,
                        uint32_t num_shells, const uint8_t *species_z, float *out) {
  return prefactors_t<float>(sm, n, shells, num_shells, species_z, out);
}
int sqsg_supercell_f32(const float *lat, const float *frac, const uint8_t *species, uint32_t n,
                       uint32_t a, uint32_t b, uint32_t c, float *lattice_out, float *frac_out,
                       uint8_t *species_out) {
  return supercell_t<float>(lat, frac, species, n, a, b, c, lattice_out, frac_out, species_out);
}

int sqsg_num_permutations(const uint8_t *conf, uint32_t n, char *out, size_t cap) {
  return guard([&] {
    if (n) need(conf, "conf");
    copy_str(sqsg::num_permutations(conf, n).str(), out, cap);
  });
}

int sqsg_rank_permutation(const uint8_t *conf, uint32_t n, char *out, size_t cap) {
  return guard([&] {
    if (n) need(conf, "conf");
    copy_str(sqsg::rank_permutation(conf, n).str(), out, cap);
  });
}

int sqsg_unrank_permutation(uint8_t *conf, uint32_t n, const char *rank) {
  return guard([&] {
    if (n) need(conf, "conf");
    need(rank, "rank");
    const sqsg::BigUint r = sqsg::BigUint::from_string(rank);
    if (!r.fits_u128()) throw std::out_of_range("The rank is larger than the total number of permutations");
    sqsg::unrank_permutation(conf, n, r.to_u128());
  });
}

int sqsg_job_create(const sqsg_config *cfg, sqsg_job **out) {
  return guard([&] {
    need(cfg, "cfg"), need(out, "out");
    *out = nullptr;
    *out = sqsg::create_job(*cfg);
  });
}

void sqsg_job_destroy(sqsg_job *job) { delete job; }

int sqsg_job_num_problems(sqsg_job *job, uint32_t *n) {
  return guard([&] {
    need(job, "job"), need(n, "num_problems");
    *n = static_cast<uint32_t>(job->problems.size());
  });
}

int sqsg_job_problem_info(sqsg_job *job, uint32_t p, sqsg_problem_info *info) {
  return guard([&] {
    need(job, "job"), need(info, "info");
    if (p >= job->problems.size()) throw std::out_of_range("problem index out of range");
    const auto &P = job->problems[p];
    std::memset(info, 0, sizeof(*info));
    info->num_sites = P.N, info->num_species = P.K, info->num_shells = P.S;
    info->num_pairs = static_cast<uint32_t>(P.pi.size());
    info->num_bounds = static_cast<uint32_t>(P.bounds.size() / 2);
    for (size_t b = 0; b < P.bounds.size() && b < 16; ++b) info->bounds[b] = P.bounds[b];
    for (size_t k = 0; k < P.zmap.size() && k < 64; ++k) info->species_z[k] = P.zmap[k];
    for (size_t s = 0; s < P.shell_ids.size() && s < 64; ++s) info->shell_ids[s] = P.shell_ids[s];
  });
}

int sqsg_job_problem_arrays(sqsg_job *job, uint32_t p, uint32_t *pairs, uint8_t *species_packed,
                            uint32_t *sort_order, double *pref, double *pw, double *target) {
  return guard([&] {
    need(job, "job");
    if (p >= job->problems.size()) throw std::out_of_range("problem index out of range");
    const auto &P = job->problems[p];
    if (pairs)
      for (size_t q = 0; q < P.pi.size(); ++q) {
        pairs[3 * q] = P.pi[q], pairs[3 * q + 1] = P.pj[q], pairs[3 * q + 2] = P.ps[q];
      }
    if (species_packed) std::memcpy(species_packed, P.species_packed.data(), P.N);
    if (sort_order) std::memcpy(sort_order, P.sort_order.data(), P.N * sizeof(uint32_t));
    if (pref) std::memcpy(pref, P.pref.data(), P.pref.size() * sizeof(double));
    if (pw) std::memcpy(pw, P.pw.data(), P.pw.size() * sizeof(double));
    if (target) std::memcpy(target, P.tgt.data(), P.tgt.size() * sizeof(double));
  });
}

int sqsg_score_batch(sqsg_job *job, uint32_t p, const uint8_t *species, uint32_t B, int32_t *bonds,
                     double *objective, double *sro, int kernel) {
  return guard([&] {
    need(job, "job"), need(species, "species"), need(objective, "objective");
    if (p >= job->problems.size()) throw std::out_of_range("problem index out of range");
    const auto &P = job->problems[p];
    for (size_t q = 0; q < size_t(B) * P.N; ++q)
      if (species[q] >= P.K) throw std::out_of_range("species index exceeds the species count");
    if (B == 0) return;
    sqsg::ensure_device(job);
    const size_t nb = size_t(P.S) * P.K * P.K;
    sqsg::DevBuf d_sp(size_t(B) * P.N), d_obj(size_t(B) * sizeof(double));
    sqsg::DevBuf d_bonds(bonds ? size_t(B) * nb * sizeof(int32_t) : 0);
    sqsg::DevBuf d_sro(sro ? size_t(B) * nb * sizeof(double) : 0);
    SQSG_HIP(hipMemcpy(d_sp.as<void>(), species, size_t(B) * P.N, hipMemcpyHostToDevice));
    sqsg::score_batch(job, p, d_sp.as<uint8_t>(), B, bonds ? d_bonds.as<int32_t>() : nullptr,
                      d_obj.as<double>(), sro ? d_sro.as<double>() : nullptr, kernel, sqsg::job_stream(job));
    SQSG_HIP(hipStreamSynchronize(sqsg::job_stream(job)));
    SQSG_HIP(hipMemcpy(objective, d_obj.as<void>(), size_t(B) * sizeof(double),
                       hipMemcpyDeviceToHost));
    if (bonds)
      SQSG_HIP(hipMemcpy(bonds, d_bonds.as<void>(), size_t(B) * nb * sizeof(int32_t),
                         hipMemcpyDeviceToHost));
    if (sro)
      SQSG_HIP(hipMemcpy(sro, d_sro.as<void>(), size_t(B) * nb * sizeof(double),
                         hipMemcpyDeviceToHost));
  });
}

int sqsg_score_batch_device(sqsg_job *job, uint32_t p, const uint8_t *d_species, uint32_t B,
                            int32_t *d_bonds, double *d_objective, double *d_sro, int kernel,
                            void *stream) {
  return guard([&] {
    need(job, "job"), need(d_species, "species"), need(d_objective, "objective");
    sqsg::score_batch(job, p, d_species, B, d_bonds, d_objective, d_sro, kernel,
                      static_cast<hipStream_t>(stream));
  });
}

int sqsg_job_set_kernel(sqsg_job *job, int kernel) {
  return guard([&] {
    need(job, "job");
    if (kernel < SQSG_KERNEL_AUTO || kernel > SQSG_KERNEL_DENSE_PAIRED)
      throw std::invalid_argument("unknown kernel variant");
    job->kernel = kernel;
  });
}

int sqsg_job_set_pruning(sqsg_job *job, int enable) {
  return guard([&] {
    need(job, "job");
    if (enable)
      for (const auto &P : job->problems)
        for (uint32_t t : P.terms) {
          const uint32_t s = t & 0xff, xi = (t >> 8) & 0xff, eta = t >> 16;
          if (!(P.pw[(s * P.K + xi) * P.K + eta] >= 0.0))
            throw std::invalid_argument("pruning needs non-negative pair weights");
        }
    job->prune = enable != 0;
  });
}

int sqsg_job_set_launch_size(sqsg_job *job, uint64_t n) {
  return guard([&] {
    need(job, "job");
    job->launch_size = n;
  });
}

int sqsg_run(sqsg_job *job, sqsg_progress_fn cb, void *user) {
  return guard([&] {
    need(job, "job");
    sqsg::run_job(job, cb, user);
  });
}

void sqsg_cancel(sqsg_job *job) {
  if (job) job->cancel.store(true);
}

int sqsg_job_statistics(sqsg_job *job, sqsg_statistics *stats) {
  return guard([&] {
    need(job, "job"), need(stats, "stats");
    *stats = job->stats;
  });
}

int sqsg_job_timing(sqsg_job *job, double *kernel_ms, uint64_t *launches, uint64_t *candidates) {
  return guard([&] {
    need(job, "job");
    if (kernel_ms) *kernel_ms = job->kernel_ms;
    if (launches) *launches = job->launches;
    if (candidates) *candidates = job->scored;
  });
}

int sqsg_result_count(sqsg_job *job, uint64_t *num_results, uint64_t *num_objectives) {
  return guard([&] {
    need(job, "job");
    if (num_results) *num_results = job->res.size();
    if (num_objectives) *num_objectives = job->res.level_obj.size();
  });
}

int sqsg_result_get(sqsg_job *job, uint64_t r, double *objective, uint8_t *species_z,
                    double *sub_objectives, double *sro, char *rank, size_t rank_cap,
                    uint64_t *candidate) {
  return guard([&] {
    need(job, "job");
    const auto &R = job->res;
    if (r >= R.size()) throw std::out_of_range("result index out of range");
    const size_t nprob = job->problems.size();
    if (objective) *objective = sqsg::result_objective(job, r);
    if (species_z) sqsg::result_species_z(job, r, species_z);
    if (sub_objectives)
      std::memcpy(sub_objectives, R.sub.data() + r * nprob, nprob * sizeof(double));
    if (sro) sqsg::result_sro(job, r, 1, sro);
    if (rank) copy_str(sqsg::result_rank(job, r), rank, rank_cap);
    if (candidate) *candidate = R.cand[r];
  });
}

int sqsg_result_species_packed(sqsg_job *job, uint64_t r, uint8_t *species) {
  return guard([&] {
    need(job, "job"), need(species, "species");
    if (r >= job->res.size()) throw std::out_of_range("result index out of range");
    std::memcpy(species, job->res.species.data() + r * job->stride, job->stride);
  });
}

int sqsg_result_levels(sqsg_job *job, double *objectives, uint64_t *level_start) {
  return guard([&] {
    need(job, "job");
    const auto &R = job->res;
    if (objectives)
      std::memcpy(objectives, R.level_obj.data(), R.level_obj.size() * sizeof(double));
    if (level_start)
      std::memcpy(level_start, R.level_start.data(), R.level_start.size() * sizeof(uint64_t));
  });
}

int sqsg_result_bulk(sqsg_job *job, uint64_t first, uint64_t count, double *objective,
                     uint64_t *candidate, uint8_t *species_z, uint64_t *rank) {
  return guard([&] {
    need(job, "job");
    sqsg::result_bulk(job, first, count, objective, candidate, species_z, rank);
  });
}

int sqsg_result_sro_bulk(sqsg_job *job, uint64_t first, uint64_t count, double *sro) {
  return guard([&] {
    need(job, "job"), need(sro, "sro");
    sqsg::result_sro(job, first, count, sro);
  });
}

int sqsg_batch_run(sqsg_job **jobs, uint32_t num_jobs, int use_graph, sqsg_batch_report *report) {
  return guard([&] {
    need(jobs, "jobs");
    for (uint32_t k = 0; k < num_jobs; ++k) need(jobs[k], "job");
    sqsg::batch_run(jobs, num_jobs, use_graph, report);
  });
}

int sqsg_job_kernel_shape(sqsg_job *job, char *buf, size_t len) {
  return guard([&] {
    need(job, "job");
    need(buf, "buf");
    if (len == 0) throw std::invalid_argument("empty buffer");
    const std::string s = sqsg::kernel_shape(job);
    const size_t n = std::min(len - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
  });
}

int sqsg_job_set_dense_launch(sqsg_job *job, uint32_t shape, uint32_t passes, uint64_t launch) {
  return guard([&] {
    need(job, "job");
    sqsg::analyze_job(job);
    if (shape && (job->problems.size() != 1 ||
                  shape > sqsg::dense_alt_shapes(sqsg::dense_tile_bound_of(job))))
      throw std::out_of_range("no such alternative dense launch shape for this job");
    job->dense_alt = shape;
    job->dense_passes = passes;
    job->tuned_launch = launch;
  });
}

int sqsg_job_dense_launch(sqsg_job *job, uint32_t *num_shapes, uint32_t *shape,
                          uint32_t *passes, uint64_t *launch) {
  return guard([&] {
    need(job, "job");
    sqsg::analyze_job(job);
    if (num_shapes)
      *num_shapes = job->dense_ok && job->problems.size() == 1
                        ? 1 + sqsg::dense_alt_shapes(sqsg::dense_tile_bound_of(job))
                        : (job->dense_ok ? 1u : 0u);
    if (shape) *shape = job->dense_alt;
    if (passes) *passes = job->dense_passes;
    if (launch) *launch = job->tuned_launch;
  });
}

int sqsg_autotune(sqsg_job *job, int *chosen, double *cand_per_s) {
  return guard([&] {
    need(job, "job");
    sqsg::autotune(job, chosen, cand_per_s);
  });
}

int sqsg_server_create(const sqsg_server_options *opt, sqsg_server **out) {
  return guard([&] {
    need(out, "out");
    sqsg_server_options o{};
    if (opt) o = *opt;
    *out = sqsg::server_create(o);
  });
}

int sqsg_server_submit(sqsg_server *srv, sqsg_job *job, uint64_t *ticket) {
  return guard([&] {
    need(srv, "server"), need(job, "job");
    const uint64_t t = sqsg::server_submit(srv, job);
    if (ticket) *ticket = t;
  });
}

int sqsg_server_step(sqsg_server *srv, int *active_jobs, int *queued_jobs) {
  return guard([&] {
    need(srv, "server");
    sqsg::server_step(srv, active_jobs, queued_jobs);
  });
}

int sqsg_server_poll(sqsg_server *srv, uint64_t ticket, int *done, double *t_submit,
                     double *t_admit, double *t_done) {
  return guard([&] {
    need(srv, "server");
    sqsg::server_poll(srv, ticket, done, t_submit, t_admit, t_done);
  });
}

void sqsg_server_destroy(sqsg_server *srv) { sqsg::server_destroy(srv); }

}  // extern "C"
