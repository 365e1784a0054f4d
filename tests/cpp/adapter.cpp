// Reference-side adapter for libsqsgen_amd: what sqsgen::run_optimization
// (include/sqsgen/sqs.h:625-634) calls for a configuration<double>, written against
// include/sqsgen_amd.h alone and compiled with a plain host compiler.
//
//   configuration       mirrors core::configuration<double> (core/config.h:32-48): structure
//                       after the supercell, composition per sublattice, per-problem shell
//                       radii and weights, and the prefactor / pair-weight / target cubes held
//                       column-major like the reference's Eigen tensors;
//   flatten()           -> sqsg_config (cubes transposed to the ABI's row-major [S][K][K]);
//   check()             -> the reference's exception types (sqs.h:218-221,620,
//                          core/results.h:172-192);
//   run_optimization_mi355x()  job create, run with a progress callback that can stop the
//                       run (sqs_callback_context::stop, types.h:222-232), pack rebuilt;
//   unflatten_results() -> result_pack: levels in ascending objective order, each a list of
//                       results with species (Z, original site order), SRO cubes and the
//                       decimal rank (sqs_result_pack, core/results.h:336-389).
//
// Modes:
//   adapter host        host-only entry points (geometry, ranks, configuration checks and
//                       their exception mapping); needs no device
//   adapter run ITER    optimises the 32-site FCC Al16Ti16 cell (3 shells, weights 1, 1/2,
//                       1/3, pair weights 1, seed 42, chunk 100000) on the GPU and prints the
//                       pack as JSON
//   adapter stop        the same run stopped from the callback after its first launch
#include <sqsgen_amd.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace adapter {

// ---- status codes -> the reference's exceptions ------------------------------------------
struct stop_requested : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void check(int status) {
  if (status == SQSG_OK) return;
  const std::string msg = sqsg_last_error();
  switch (status) {
    case SQSG_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case SQSG_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
    case SQSG_ERR_DOMAIN: throw std::domain_error(msg);
    case SQSG_ERR_NO_MEMORY: throw std::bad_alloc();
    case SQSG_ERR_CANCELLED: throw stop_requested(msg);
    default: throw std::runtime_error(msg);
  }
}

// ---- configuration<double> ---------------------------------------------------------------
// A column-major S x K x K tensor as Eigen::Tensor<double, 3, ColMajor> stores it: element
// (s, a, b) at s + S * (a + K * b).
struct cube {
  uint32_t S = 0, K = 0;
  std::vector<double> data;
  cube() = default;
  cube(uint32_t s, uint32_t k, double fill) : S(s), K(k), data(size_t(s) * k * k, fill) {}
  double &operator()(uint32_t s, uint32_t a, uint32_t b) { return data[s + S * (a + K * b)]; }
  double operator()(uint32_t s, uint32_t a, uint32_t b) const {
    return data[s + S * (a + K * b)];
  }
};

struct sublattice {
  std::vector<uint32_t> sites;
  std::map<uint8_t, uint32_t> composition;  // Z -> amount, ascending Z
};

struct configuration {
  bool split = false, systematic = false;
  std::vector<double> lattice;      // 9, rows a, b, c
  std::vector<double> frac_coords;  // N x 3
  std::vector<uint8_t> species;     // N, Z
  std::vector<sublattice> composition;
  std::vector<std::vector<double>> shell_radii;                // per problem
  std::vector<std::map<uint32_t, double>> shell_weights;       // per problem
  std::vector<cube> prefactors, pair_weights, target_objective;  // per problem
  uint64_t iterations = 0, chunk_size = 0, keep = 1;
  std::vector<std::pair<bool, uint64_t>> seeds;  // per problem: present, value
  int64_t max_results_per_objective = -1;
};

// Owns the flattened arrays an sqsg_config points into.
struct flat {
  std::vector<uint32_t> site_off, sites, comp_off, comp_amounts, radii_off, shell_off, shells, K;
  std::vector<uint8_t> comp_species, seed_present;
  std::vector<double> radii, weights, pref, pw, tgt;
  std::vector<uint64_t> seeds;
  sqsg_config c{};
};

void append_row_major(std::vector<double> &out, const cube &t) {
  for (uint32_t s = 0; s < t.S; ++s)
    for (uint32_t a = 0; a < t.K; ++a)
      for (uint32_t b = 0; b < t.K; ++b) out.push_back(t(s, a, b));
}

void flatten(const configuration &cfg, flat &f) {
  f.site_off = {0};
  f.comp_off = {0};
  for (const auto &sl : cfg.composition) {
    f.sites.insert(f.sites.end(), sl.sites.begin(), sl.sites.end());
    f.site_off.push_back(static_cast<uint32_t>(f.sites.size()));
    for (const auto &[z, n] : sl.composition) {
      f.comp_species.push_back(z);
      f.comp_amounts.push_back(n);
    }
    f.comp_off.push_back(static_cast<uint32_t>(f.comp_species.size()));
  }
  f.radii_off = {0};
  f.shell_off = {0};
  for (size_t q = 0; q < cfg.shell_radii.size(); ++q) {
    f.radii.insert(f.radii.end(), cfg.shell_radii[q].begin(), cfg.shell_radii[q].end());
    f.radii_off.push_back(static_cast<uint32_t>(f.radii.size()));
    for (const auto &[s, w] : cfg.shell_weights[q]) {
      f.shells.push_back(s);
      f.weights.push_back(w);
    }
    f.shell_off.push_back(static_cast<uint32_t>(f.shells.size()));
    f.K.push_back(cfg.prefactors[q].K);
    append_row_major(f.pref, cfg.prefactors[q]);
    append_row_major(f.pw, cfg.pair_weights[q]);
    append_row_major(f.tgt, cfg.target_objective[q]);
    f.seed_present.push_back(cfg.seeds[q].first ? 1 : 0);
    f.seeds.push_back(cfg.seeds[q].second);
  }
  sqsg_config &c = f.c;
  c.num_sites = static_cast<uint32_t>(cfg.species.size());
  c.lattice = cfg.lattice.data();
  c.frac_coords = cfg.frac_coords.data();
  c.species = cfg.species.data();
  c.sublattice_mode = cfg.split ? SQSG_SUBLATTICE_SPLIT : SQSG_SUBLATTICE_INTERACT;
  c.iteration_mode = cfg.systematic ? SQSG_ITERATION_SYSTEMATIC : SQSG_ITERATION_RANDOM;
  c.num_sublattices = static_cast<uint32_t>(cfg.composition.size());
  c.site_offsets = f.site_off.data();
  c.sites = f.sites.data();
  c.comp_offsets = f.comp_off.data();
  c.comp_species = f.comp_species.data();
  c.comp_amounts = f.comp_amounts.data();
  c.radii_offsets = f.radii_off.data();
  c.radii = f.radii.data();
  c.shell_offsets = f.shell_off.data();
  c.shells = f.shells.data();
  c.shell_weights = f.weights.data();
  c.num_species = f.K.data();
  c.prefactors = f.pref.data();
  c.pair_weights = f.pw.data();
  c.target_objective = f.tgt.data();
  c.iterations = cfg.iterations;
  c.chunk_size = cfg.chunk_size;
  c.seeds = f.seeds.data();
  c.seed_present = f.seed_present.data();
  c.keep = cfg.keep;
  c.max_results_per_objective = cfg.max_results_per_objective;
}

// ---- result pack ---------------------------------------------------------------------------
struct result {
  double objective;
  std::vector<uint8_t> species;  // Z, original site order
  std::vector<double> sro;       // concatenated [S][K][K] per problem
  std::string rank;
};
using result_pack = std::vector<std::pair<double, std::vector<result>>>;

struct job_handle {
  sqsg_job *job = nullptr;
  ~job_handle() { sqsg_job_destroy(job); }
};

result_pack unflatten_results(sqsg_job *job, const configuration &cfg) {
  uint64_t n = 0, levels = 0;
  check(sqsg_result_count(job, &n, &levels));
  std::vector<double> objs(levels);
  std::vector<uint64_t> start(levels + 1);
  check(sqsg_result_levels(job, objs.data(), start.data()));
  size_t cube_len = 0;
  for (const auto &p : cfg.prefactors) cube_len += p.data.size();
  const size_t nprob = cfg.prefactors.size();
  result_pack pack;
  for (uint64_t l = 0; l < levels; ++l) {
    std::vector<result> rs;
    for (uint64_t r = start[l]; r < start[l + 1]; ++r) {
      result x;
      x.species.resize(cfg.species.size());
      x.sro.resize(cube_len);
      std::vector<double> subs(nprob);
      char rank[128];
      uint64_t cand = 0;
      check(sqsg_result_get(job, r, &x.objective, x.species.data(), subs.data(), x.sro.data(),
                            rank, sizeof rank, &cand));
      x.rank = rank;
      rs.push_back(std::move(x));
    }
    pack.emplace_back(objs[l], std::move(rs));
  }
  return pack;
}

struct progress {
  int launches = 0, stop_after = -1;
  uint64_t finished = 0;
};

int on_progress(const sqsg_statistics *st, void *user) {
  auto *p = static_cast<progress *>(user);
  ++p->launches;
  p->finished = st->finished;
  return p->stop_after > 0 && p->launches >= p->stop_after;  // non-zero: ctx.stop()
}

result_pack run_optimization_mi355x(const configuration &cfg, progress *prog,
                                    uint64_t launch_size, uint64_t *finished) {
  flat f;
  flatten(cfg, f);
  job_handle h;
  check(sqsg_job_create(&f.c, &h.job));
  if (launch_size) check(sqsg_job_set_launch_size(h.job, launch_size));
  const int st = sqsg_run(h.job, &on_progress, prog);
  if (st != SQSG_ERR_CANCELLED) check(st);  // a stopped run still returns its partial pack
  sqsg_statistics stats{};
  check(sqsg_job_statistics(h.job, &stats));
  *finished = stats.finished;
  return unflatten_results(h.job, cfg);
}

// ---- the 32-site workload: parse steps of the reference done with the host entry points ----
configuration fcc32(uint64_t iterations) {
  const double a = 4.05;
  const double lat[9] = {a, 0, 0, 0, a, 0, 0, 0, a};
  const double basis[12] = {0, 0, 0, 0.5, 0.5, 0, 0.5, 0, 0.5, 0, 0.5, 0.5};
  const uint8_t al[4] = {13, 13, 13, 13};
  configuration cfg;
  cfg.lattice.resize(9);
  cfg.frac_coords.resize(32 * 3);
  cfg.species.resize(32);
  check(sqsg_supercell(lat, basis, al, 4, 2, 2, 2, cfg.lattice.data(), cfg.frac_coords.data(),
                       cfg.species.data()));
  sublattice sl;
  for (uint32_t i = 0; i < 32; ++i) sl.sites.push_back(i);
  sl.composition = {{13, 16}, {22, 16}};
  cfg.composition = {sl};
  // the structure with the composition applied (structure.h:390-399)
  std::vector<uint8_t> applied(32);
  for (uint32_t i = 0; i < 32; ++i) applied[i] = i < 16 ? 13 : 22;
  std::vector<double> dm(32 * 32);
  check(sqsg_distance_matrix(cfg.lattice.data(), cfg.frac_coords.data(), 32, dm.data()));
  double radii[64];
  uint32_t nr = 0;
  check(sqsg_distances_naive(dm.data(), 32, 1e-3, 1e-5, radii, 64, &nr));  // shell_radii.h:19-20
  cfg.shell_radii = {std::vector<double>(radii, radii + nr)};
  cfg.shell_weights = {{{1, 1.0}, {2, 0.5}, {3, 1.0 / 3.0}}};
  std::vector<uint32_t> sm(32 * 32);
  const double eps = 2.220446049250313e-16;
  check(sqsg_shell_matrix(dm.data(), 32, radii, nr, eps, 1e-9, sm.data()));
  const uint32_t shells[3] = {1, 2, 3};
  std::vector<double> pref(3 * 2 * 2);
  check(sqsg_prefactors(sm.data(), 32, shells, 3, applied.data(), pref.data()));
  cube p(3, 2, 0.0);
  for (uint32_t s = 0; s < 3; ++s)
    for (uint32_t x = 0; x < 2; ++x)
      for (uint32_t y = 0; y < 2; ++y) p(s, x, y) = pref[(s * 2 + x) * 2 + y];
  cfg.prefactors = {p};
  cfg.pair_weights = {cube(3, 2, 1.0)};
  cfg.target_objective = {cube(3, 2, 0.0)};
  cfg.iterations = iterations;
  cfg.chunk_size = 100000 < iterations ? 100000 : iterations;
  cfg.seeds = {{true, 42}};
  return cfg;
}

void print_pack(const result_pack &pack, uint64_t finished, const progress &prog) {
  std::printf("{\"finished\": %llu, \"launches\": %d, \"levels\": [",
              static_cast<unsigned long long>(finished), prog.launches);
  for (size_t l = 0; l < pack.size(); ++l) {
    std::printf("%s{\"objective\": \"%a\", \"results\": [", l ? ", " : "", pack[l].first);
    for (size_t r = 0; r < pack[l].second.size(); ++r) {
      const result &x = pack[l].second[r];
      std::printf("%s{\"objective\": \"%a\", \"rank\": \"%s\", \"species\": [", r ? ", " : "",
                  x.objective, x.rank.c_str());
      for (size_t i = 0; i < x.species.size(); ++i) std::printf("%s%u", i ? ", " : "", x.species[i]);
      std::printf("], \"sro\": [");
      for (size_t i = 0; i < x.sro.size(); ++i) std::printf("%s\"%a\"", i ? ", " : "", x.sro[i]);
      std::printf("]}");
    }
    std::printf("]}");
  }
  std::printf("]}\n");
}

template <class E, class F> bool throws(F &&f) {
  try {
    f();
  } catch (const E &) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}

int host_checks() {
  int failures = 0;
  auto expect = [&](bool ok, const char *what) {
    if (!ok) {
      std::fprintf(stderr, "FAILED: %s\n", what);
      ++failures;
    }
  };
  expect(sqsg_version() > 0, "version");
  configuration cfg = fcc32(1000);
  // naive radii: 0, then the FCC shells a / sqrt(2), a, a sqrt(3 / 2)
  expect(cfg.shell_radii[0].size() >= 4, "three shells");
  const double want[4] = {0.0, 4.05 / std::sqrt(2.0), 4.05, 4.05 * std::sqrt(1.5)};
  for (int s = 0; s < 4; ++s)
    expect(std::fabs(cfg.shell_radii[0][s] - want[s]) < 1e-9, "FCC shell radii");
  // prefactors of an A16B16 cell: 1 / (N M_s x_a x_b) with x = 1/2 and the minimum-image
  // coordination numbers of the 2 x 2 x 2 cell (8.1 A box): 12, 3 (the +-a neighbours of the
  // second shell are one site) and 12 (the 24 third neighbours pair up)
  const double m[3] = {12, 3, 12};
  for (uint32_t s = 0; s < 3; ++s)
    expect(std::fabs(cfg.prefactors[0](s, 0, 1) - 1.0 / (32 * m[s] * 0.25)) < 1e-12,
           "prefactor 1 / (N M_s x_a x_b)");
  // ranks of the first and last multiset permutations of 0^16 1^16
  std::vector<uint8_t> conf(32);
  for (uint32_t i = 0; i < 32; ++i) conf[i] = i < 16 ? 0 : 1;
  char buf[64];
  check(sqsg_rank_permutation(conf.data(), 32, buf, sizeof buf));
  expect(std::string(buf) == "1", "rank of the sorted configuration");
  for (uint32_t i = 0; i < 32; ++i) conf[i] = i < 16 ? 1 : 0;
  check(sqsg_rank_permutation(conf.data(), 32, buf, sizeof buf));
  expect(std::string(buf) == "601080390", "rank of the last configuration = C(32, 16)");
  check(sqsg_num_permutations(conf.data(), 32, buf, sizeof buf));
  expect(std::string(buf) == "601080390", "number of permutations");
  // job creation is host-side: a valid configuration builds its problem, invalid ones map
  // onto the reference's exception types
  {
    flat f;
    flatten(cfg, f);
    job_handle h;
    check(sqsg_job_create(&f.c, &h.job));
    sqsg_problem_info info{};
    check(sqsg_job_problem_info(h.job, 0, &info));
    expect(info.num_sites == 32 && info.num_species == 2 && info.num_shells == 3, "problem");
    expect(info.num_pairs == 32 * (12 + 3 + 12) / 2, "pairs of the three shells");
  }
  expect(throws<std::out_of_range>([&] {
           configuration bad = cfg;
           bad.composition[0].composition = {{13, 16}, {22, 15}};  // 31 of 32 sites
           flat f;
           flatten(bad, f);
           job_handle h;
           check(sqsg_job_create(&f.c, &h.job));
         }),
         "composition not matching the sublattice -> std::out_of_range");
  expect(throws<std::invalid_argument>([&] {
           flat f;
           flatten(cfg, f);
           f.c.num_sites = 0;
           job_handle h;
           check(sqsg_job_create(&f.c, &h.job));
         }) || throws<std::out_of_range>([&] {
           flat f;
           flatten(cfg, f);
           f.c.num_sites = 0;
           job_handle h;
           check(sqsg_job_create(&f.c, &h.job));
         }),
         "empty structure rejected");
  expect(throws<std::invalid_argument>([] { check(sqsg_job_create(nullptr, nullptr)); }),
         "null arguments -> std::invalid_argument");
  std::printf("{\"host_checks_failed\": %d}\n", failures);
  return failures ? 1 : 0;
}

}  // namespace adapter

int main(int argc, char **argv) {
  using namespace adapter;
  const std::string mode = argc > 1 ? argv[1] : "host";
  try {
    if (mode == "host") return host_checks();
    const uint64_t iterations = argc > 2 ? std::stoull(argv[2]) : 200000;
    configuration cfg = fcc32(iterations);
    progress prog;
    if (mode == "stop") prog.stop_after = 1;
    uint64_t finished = 0;
    // launches of 50000 candidates, so that a stop after the first leaves a partial run
    const result_pack pack = run_optimization_mi355x(cfg, &prog, 50000, &finished);
    print_pack(pack, finished, prog);
    return 0;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "adapter: %s\n", e.what());
    return 2;
  }
}
