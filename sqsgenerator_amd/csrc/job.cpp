// Job construction (optimization_config<T,Mode>::from_config), the run driver (the
// worker loop of optimizer::run, sqs.h:299-580, moved onto the device) and result
// assembly (sqs_result_collection / postprocess_results / sqs_result_wrapper::rank).
#include "job.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <random>
#include <set>

#include "kernels.hpp"
#include "permutation.hpp"

namespace sqsg {

const char *const kDenseDomain =
    "the dense kernel needs problems of at most 1024 sites, 16 species and 127 upper "
    "neighbours per site and shell (split jobs: Monte Carlo, at most 256 sites per "
    "sublattice) and a configuration in double precision";

void hip_check(hipError_t e, const char *what) {
  if (e == hipErrorOutOfMemory) throw std::bad_alloc();
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

void DevBuf::alloc(size_t bytes) {
  reset();
  if (bytes == 0) bytes = 16;
  void *p = nullptr;
  hip_check(hipMalloc(&p, bytes), "hipMalloc");
  p_ = p;
  n_ = bytes;
}

void DevBuf::reset() {
  if (p_) (void)hipFree(p_);
  p_ = nullptr;
  n_ = 0;
}

int device_cu_count() {
  static int cus = [] {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    return prop.multiProcessorCount;
  }();
  return cus;
}

namespace {

template <class T> void upload(DevBuf &b, const std::vector<T> &v, size_t pad_to = 1) {
  const size_t bytes = v.size() * sizeof(T);
  const size_t padded = ((bytes + pad_to - 1) / pad_to) * pad_to;
  b.alloc(padded);
  if (padded != bytes) SQSG_HIP(hipMemset(b.as<void>(), 0, padded));
  if (bytes) SQSG_HIP(hipMemcpy(b.as<void>(), v.data(), bytes, hipMemcpyHostToDevice));
}

uint64_t make_random_seed() {  // shuffle.h:16-21
  std::random_device rd;
  std::mt19937_64 gen(rd());
  std::uniform_int_distribution<uint64_t> dis;
  return dis(gen);
}

struct ProblemInput {
  std::vector<double> frac;          // n x 3
  std::vector<uint8_t> z;            // n
  std::vector<uint32_t> sort_order;  // sorted position -> input site
  std::vector<uint32_t> bounds;
  std::vector<uint32_t> sites;       // input site -> original structure site
};

void build_problem(Problem &P, const ProblemInput &in, const Lattice &lat, const sqsg_config &c,
                   uint32_t p, bool systematic) {
  const uint32_t n = static_cast<uint32_t>(in.z.size());
  if (n == 0) throw std::invalid_argument("Cannot create a structure without atoms");
  if (n > kMaxSites)
    throw std::out_of_range("structures of more than 4096 sites are not supported on device");
  P.N = n;
  P.sort_order = in.sort_order;
  P.bounds = in.bounds;
  P.sites = in.sites;
  // species map over the sorted structure (optimization_config.h:129-133)
  std::set<uint8_t> zs(in.z.begin(), in.z.end());
  P.zmap.assign(zs.begin(), zs.end());
  P.K = static_cast<uint32_t>(P.zmap.size());
  if (P.K != c.num_species[p])
    throw std::invalid_argument("Incompatible configuration: species count does not match the "
                                "extent of the parameter arrays");
  uint8_t zidx[256];
  for (uint32_t k = 0; k < P.K; ++k) zidx[P.zmap[k]] = static_cast<uint8_t>(k);
  P.species_packed.resize(n);
  for (uint32_t i = 0; i < n; ++i) P.species_packed[i] = zidx[in.z[i]];
  // shells and weights (shell_weights_t is an ordered map)
  const uint32_t s0 = c.shell_offsets[p], s1 = c.shell_offsets[p + 1];
  if (s1 <= s0) throw std::out_of_range("no coordination shells selected");
  P.S = s1 - s0;
  if (P.S > 255) throw std::out_of_range("at most 255 coordination shells are supported");
  P.shell_ids.assign(c.shells + s0, c.shells + s1);
  for (uint32_t s = 1; s < P.S; ++s)
    if (P.shell_ids[s] <= P.shell_ids[s - 1])
      throw std::invalid_argument("shell indices must be strictly ascending");
  std::vector<double> weights(c.shell_weights + s0, c.shell_weights + s1);
  std::vector<double> radii(c.radii + c.radii_offsets[p], c.radii + c.radii_offsets[p + 1]);
  // pair list of the sorted structure (structure.h:434-451; optimization_config.h:134)
  std::vector<uint32_t> sm;
  if (c.precision == SQSG_PREC_SINGLE) {  // structure<float>: the geometry in float
    LatticeT<float> latf;
    for (int k = 0; k < 9; ++k) latf.m[k] = static_cast<float>(lat.m[k]);
    const std::vector<float> fracf(in.frac.begin(), in.frac.end());
    const std::vector<float> radf(radii.begin(), radii.end());
    const std::vector<float> dm = distance_matrix<float>(latf, fracf.data(), n);
    sm = shell_matrix<float>(dm, n, radf, kEpsF, 1.0e-9f);
  } else {
    const std::vector<double> dm = distance_matrix<double>(lat, in.frac.data(), n);
    sm = shell_matrix<double>(dm, n, radii, kEps, 1.0e-9);
  }
  std::map<uint32_t, uint32_t> shell_index;
  for (uint32_t s = 0; s < P.S; ++s) shell_index[P.shell_ids[s]] = s;
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j = i + 1; j < n; ++j) {
      auto it = shell_index.find(sm[size_t(i) * n + j]);
      if (it == shell_index.end()) continue;
      P.pi.push_back(i);
      P.pj.push_back(j);
      P.ps.push_back(it->second);
    }
  // parameter cubes [S][K][K]; pair weights scaled per shell (optimization.h:17-33)
  size_t cube_off = 0;
  for (uint32_t q = 0; q < p; ++q) {
    const size_t Sq = c.shell_offsets[q + 1] - c.shell_offsets[q];
    cube_off += Sq * c.num_species[q] * c.num_species[q];
  }
  const size_t cube = size_t(P.S) * P.K * P.K;
  P.pref.assign(c.prefactors + cube_off, c.prefactors + cube_off + cube);
  P.pw.assign(c.pair_weights + cube_off, c.pair_weights + cube_off + cube);
  P.tgt.assign(c.target_objective + cube_off, c.target_objective + cube_off + cube);
  const uint32_t K = P.K;
  P.single = c.precision == SQSG_PREC_SINGLE;
  // w *= w_s in the configuration's type (a float product is the double product rounded)
  auto scale = [&](double x, double w) {
    return P.single ? static_cast<double>(static_cast<float>(x) * static_cast<float>(w)) : x * w;
  };
  for (uint32_t s = 0; s < P.S; ++s)
    for (uint32_t xi = 0; xi < K; ++xi)
      for (uint32_t eta = xi; eta < K; ++eta) {
        P.pw[(s * K + xi) * K + eta] = scale(P.pw[(s * K + xi) * K + eta], weights[s]);
        if (xi == eta) continue;
        P.pw[(s * K + eta) * K + xi] = scale(P.pw[(s * K + eta) * K + xi], weights[s]);
      }
  // objective terms in the order of optimization.h:85-100
  for (uint32_t s = 0; s < P.S; ++s)
    for (uint32_t xi = 0; xi < K; ++xi)
      for (uint32_t eta = xi; eta < K; ++eta) P.terms.push_back(s | (xi << 8) | (eta << 16));
  P.seed = c.seed_present && c.seed_present[p] ? c.seeds[p] : make_random_seed();
  P.draws = 0;
  for (size_t b = 0; b + 1 < P.bounds.size(); b += 2) {
    if (P.bounds[b + 1] <= P.bounds[b]) throw std::invalid_argument("empty shuffle window");
    P.draws += P.bounds[b + 1] - P.bounds[b] - 1;
  }
  if (P.bounds.size() / 2 > static_cast<size_t>(kMaxBounds))
    throw std::out_of_range("at most 8 sublattices are supported in interact mode");
  if (systematic) {
    const uint32_t lo = P.bounds[0], hi = P.bounds[1];
    std::set<uint8_t> local(P.species_packed.begin() + lo, P.species_packed.begin() + hi);
    if (local.size() > static_cast<size_t>(kMaxLocalSpecies))
      throw std::out_of_range("at most 16 species per systematic window");
    P.sys_nspecies = 0;
    for (uint8_t v : local) {
      P.sys_rev[P.sys_nspecies] = v;
      P.sys_hist[P.sys_nspecies] = static_cast<uint32_t>(
          std::count(P.species_packed.begin() + lo, P.species_packed.begin() + hi, v));
      ++P.sys_nspecies;
    }
  }
}

}  // namespace

void Problem::upload_device() {
  std::vector<uint32_t> packed(pi.size());
  for (size_t q = 0; q < packed.size(); ++q)
    packed[q] = pi[q] | (pj[q] << kPairIndexBits) | (ps[q] << 24);
  upload(d_pairs, packed);
  upload(d_species0, species_packed, 16);
  upload(d_pref, pref);
  upload(d_pw, pw);
  upload(d_tgt, tgt);
  upload(d_terms, terms);
  std::vector<TermParam> tp(terms.size());
  for (size_t t = 0; t < terms.size(); ++t) {
    const uint32_t s = terms[t] & 0xff, xi = (terms[t] >> 8) & 0xff, eta = terms[t] >> 16;
    const size_t ix = (size_t(s) * K + xi) * K + eta;
    tp[t] = TermParam{pref[ix], pw[ix], tgt[ix], xi, eta};
  }
  upload(d_tparams, tp);
}

void Problem::build_lane(uint32_t field_bits) {
  F = field_bits;
  const uint32_t cap = (1u << F) - 1u;  // real neighbours per group (F-bit fields)
  const uint32_t dummy = N << 6;          // byte offset of the dummy site's row
  // upper neighbours per (shell, site), ascending j, in groups of at most `cap`; a group is
  // a header block {i * 64, data blocks} followed by 4-entry blocks of j * 64 padded with
  // the dummy site
  std::vector<std::vector<uint32_t>> nb(size_t(S) * N);
  for (size_t q = 0; q < pi.size(); ++q) nb[size_t(ps[q]) * N + pi[q]].push_back(pj[q]);
  lane_stream.clear();
  lane_off.assign(S + 1, 0);
  for (uint32_t s = 0; s < S; ++s) {
    lane_off[s] = static_cast<uint32_t>(lane_stream.size() / 4);
    for (uint32_t i = 0; i < N; ++i) {
      const auto &js = nb[size_t(s) * N + i];
      for (size_t k = 0; k < js.size(); k += cap) {
        const uint32_t cnt = static_cast<uint32_t>(std::min<size_t>(cap, js.size() - k));
        const uint32_t blocks = (cnt + 3) / 4;
        lane_stream.insert(lane_stream.end(), {i << 6, blocks, 0u, 0u});
        for (uint32_t e = 0; e < 4 * blocks; ++e)
          lane_stream.push_back(e < cnt ? (js[k + e] << 6) : dummy);
      }
    }
  }
  lane_off[S] = static_cast<uint32_t>(lane_stream.size() / 4);
  std::vector<uint8_t> sF(species_packed.size());
  for (size_t i = 0; i < sF.size(); ++i) sF[i] = static_cast<uint8_t>(species_packed[i] * F);
  upload(d_lane_stream, lane_stream, 16);
  upload(d_lane_off, lane_off);
  upload(d_species0F, sF, 16);
}

// The last species needs no one-hot column in the dense search kernels when the
// neighbour-count sum of each species' sites is the same for every candidate.  Candidates
// permute species within the shuffle windows, so it is when every window's sites have equal
// neighbour counts in each shell (sites outside the windows keep their species); the sums
// are then those of the sorted configuration.  Problems of 2-6 species.
void Problem::reduce_columns() {
  Kc = K;
  dsum.clear();
  if (K < 2 || K > 6) return;
  std::vector<uint32_t> deg(size_t(S) * N, 0);
  for (size_t q = 0; q < pi.size(); ++q) {
    ++deg[size_t(ps[q]) * N + pi[q]];
    ++deg[size_t(ps[q]) * N + pj[q]];
  }
  for (uint32_t s = 0; s < S; ++s)
    for (size_t w = 0; w + 1 < bounds.size(); w += 2)
      for (uint32_t i = bounds[w]; i < bounds[w + 1]; ++i)
        if (deg[size_t(s) * N + i] != deg[size_t(s) * N + bounds[w]]) return;
  dsum.assign(size_t(S) * K, 0);
  for (uint32_t s = 0; s < S; ++s)
    for (uint32_t i = 0; i < N; ++i)
      dsum[size_t(s) * K + species_packed[i]] += deg[size_t(s) * N + i];
  Kc = K - 1;
}

// Tiles in the A-operand layout of v_mfma_i32_32x32x32_i8: lane l, byte q of tile (s, I, J)
// is U_s[32I + perm(l % 32)][32J + 16(l / 32) + q] with perm(m) = 16((m / 4) % 2) +
// 4(m / 8) + m % 4 (the accumulator row order of the 32 x 32 result).  Layout invariant
// the kernels rely on (their tile loads take the offset in the buffer instruction's scalar
// operand, which is not range-checked): for every shell, and for one zero shell after the
// last, the whole upper grid I <= J < TT of the kernel's tile bound TT, in (shell, I, J)
// order, zero tiles included (tiles past the problem's T tiles are zeros), so a tile's
// offset is a function of the shell and the grid position alone and a prefetch of up to
// one shell past the end stays inside the stream: dtiles_n = (S + 1) TT (TT + 1) / 2.
void Problem::build_dense(uint32_t TT) {
  if (d_dtiles.size()) return;
  if (Kc < K) upload(d_dsum, dsum);
  dT = (N + 31) / 32;
  const uint32_t n32 = 32 * dT;
  if (TT < dT) throw std::logic_error("tile bound below the problem's tile count");
  std::vector<uint8_t> sh(size_t(n32) * n32, 0xff);
  for (size_t q = 0; q < pi.size(); ++q) sh[size_t(pi[q]) * n32 + pj[q]] = static_cast<uint8_t>(ps[q]);
  dtiles_nonzero = 0;
  std::vector<uint8_t> tiles;
  for (uint32_t s = 0; s < S; ++s)
    for (uint32_t I = 0; I < TT; ++I)
      for (uint32_t J = I; J < TT; ++J) {
        const size_t base = tiles.size();
        tiles.resize(base + 1024, 0);
        if (I >= dT || J >= dT) continue;
        bool any = false;
        for (uint32_t i = 32 * I; i < 32 * I + 32 && !any; ++i)
          for (uint32_t j = 32 * J; j < 32 * J + 32; ++j)
            if (sh[size_t(i) * n32 + j] == s) {
              any = true;
              break;
            }
        if (!any) continue;
        ++dtiles_nonzero;
        for (uint32_t l = 0; l < 64; ++l) {
          const uint32_t m = l % 32;
          const uint32_t i = 32 * I + 16 * ((m / 4) % 2) + 4 * (m / 8) + m % 4;
          for (uint32_t q = 0; q < 16; ++q) {
            const uint32_t j = 32 * J + 16 * (l / 32) + q;
            tiles[base + 16 * l + q] = sh[size_t(i) * n32 + j] == s ? 1 : 0;
          }
        }
      }
  // one zero shell after the last, so that the kernels' prefetch past the last shell stays
  // inside the stream (those loads take their offset unchecked)
  tiles.resize(tiles.size() + size_t(TT) * (TT + 1) / 2 * 1024, 0);
  dtiles_n = static_cast<uint32_t>(tiles.size() / 1024);
  if (size_t(dtiles_n) != size_t(S + 1) * TT * (TT + 1) / 2)
    throw std::logic_error("dense tile stream does not hold the whole grid of every shell");
  std::vector<uint8_t> sp(n32, 0xff);
  std::copy(species_packed.begin(), species_packed.end(), sp.begin());
  upload(d_dtiles, tiles, 16);
  upload(d_species0D, sp, 16);
}

void ensure_dense(sqsg_job *job) {
  ensure_device(job);
  if (!job->dense_ok)
    throw std::invalid_argument(kDenseDomain);
  // the problems of a split job share the kernel's tile bound
  const uint32_t TT = dense_tile_bound_of(job);
  for (auto &P : job->problems) P.build_dense(TT);
}

uint32_t dense_tile_bound_of(const sqsg_job *job) {
  uint32_t tiles = 0;
  for (const auto &P : job->problems) tiles = std::max(tiles, (P.N + 31) / 32);
  return dense_tile_bound(tiles);
}

// kernel applicability (host data only)
void analyze_job(sqsg_job *job) {
  if (job->analyzed) return;
  job->kmax = 0;
  for (const auto &P : job->problems) job->kmax = std::max(job->kmax, P.K);
  job->lane_ok = job->kmax <= 8;
  for (const auto &P : job->problems) {
    if (P.N >= 65536) job->lane_ok = false;
    std::vector<size_t> per_shell(P.S, 0);
    for (uint32_t s : P.ps) ++per_shell[s];
    for (size_t c : per_shell)
      if (c >= 65536) job->lane_ok = false;  // 16-bit histogram fields
  }
  for (auto &P : job->problems) {
    P.reduce_columns();
    std::vector<uint32_t> deg(size_t(P.S) * P.N, 0);
    P.max_upper = 0;
    for (size_t q = 0; q < P.pi.size(); ++q)
      P.max_upper = std::max(P.max_upper, ++deg[size_t(P.ps[q]) * P.N + P.pi[q]]);
  }
  // one problem of up to 1024 sites, or the sublattice problems of a split job (a form for
  // Monte Carlo runs up to the tile bound dense_split_available reports: 256 sites each)
  job->dense_ok = !job->single;  // the dense kernels evaluate FP64 objectives only
  for (const auto &P : job->problems)
    job->dense_ok = job->dense_ok && P.N <= 1024 && P.K <= 16 && P.max_upper <= 127;
  if (job->problems.size() > 1)
    job->dense_ok = job->dense_ok && job->iteration_mode != SQSG_ITERATION_SYSTEMATIC &&
                    dense_split_available(dense_tile_bound_of(job));
  job->analyzed = true;
}

void ensure_device(sqsg_job *job) {
  if (job->on_device) return;
  analyze_job(job);
  for (auto &P : job->problems) P.upload_device();
  if (job->lane_ok)
    for (auto &P : job->problems) P.build_lane(lane_field_bits(job->kmax));
  job->on_device = true;
}

// The job's own stream and timing events, created on first use: jobs that only run in the
// batched or served form never hold one (a process has few hardware queues to share).
hipStream_t job_stream(sqsg_job *job) {
  if (!job->stream) {
    SQSG_HIP(hipStreamCreateWithFlags(&job->stream, hipStreamNonBlocking));
    SQSG_HIP(hipEventCreate(&job->ev0));
    SQSG_HIP(hipEventCreate(&job->ev1));
  }
  return job->stream;
}

DevProblem Problem::dev() const {
  DevProblem d{};
  d.pairs = d_pairs.as<const uint32_t>();
  d.species0 = d_species0.as<const uint8_t>();
  d.pref = d_pref.as<const double>();
  d.pw = d_pw.as<const double>();
  d.tgt = d_tgt.as<const double>();
  d.terms = d_terms.as<const uint32_t>();
  d.tparams = d_tparams.as<const TermParam>();
  d.single = single ? 1u : 0u;
  d.lane_stream = d_lane_stream.as<const uint32_t>();
  d.lane_off = d_lane_off.as<const uint32_t>();
  d.species0F = d_species0F.as<const uint8_t>();
  d.F = F;
  d.dtiles = d_dtiles.as<const uint32_t>();
  d.species0D = d_species0D.as<const uint8_t>();
  d.dsum = d_dsum.as<const uint32_t>();
  d.Kc = Kc ? Kc : K;
  d.dT = dT;
  d.dtiles_n = dtiles_n;
  d.N = N, d.K = K, d.S = S;
  d.P = static_cast<uint32_t>(pi.size());
  d.T = static_cast<uint32_t>(terms.size());
  d.nbounds = static_cast<uint32_t>(bounds.size() / 2);
  for (size_t b = 0; b < bounds.size(); ++b) d.bounds[b] = bounds[b];
  d.draws = draws;
  d.row_stride = row_stride();
  d.seed = seed;
  d.sys_nspecies = sys_nspecies;
  for (int j = 0; j < kMaxLocalSpecies; ++j) d.sys_hist[j] = sys_hist[j], d.sys_rev[j] = sys_rev[j];
  return d;
}

sqsg_job *create_job(const sqsg_config &c) {
  if (c.num_sites == 0 || !c.lattice || !c.frac_coords || !c.species)
    throw std::invalid_argument("Cannot create a structure without atoms");
  const bool systematic = c.iteration_mode == SQSG_ITERATION_SYSTEMATIC;
  const bool split = c.sublattice_mode == SQSG_SUBLATTICE_SPLIT;
  if ((c.iteration_mode != SQSG_ITERATION_RANDOM && !systematic) ||
      (c.sublattice_mode != SQSG_SUBLATTICE_INTERACT && !split) || (systematic && split))
    throw std::runtime_error("Invalid configuration of iteration and sublattice mode");
  const uint32_t N = c.num_sites, nsub = c.num_sublattices;
  if (nsub == 0) throw std::invalid_argument("You need to specify a composition");
  if (systematic && nsub != 1)
    throw std::invalid_argument("In \"systematic\" iteration mode only one sublattice is allowed");
  if (c.keep == 0)
    throw std::invalid_argument("The number of best kept structures must be greater than 0");
  if (c.iterations == 0) throw std::invalid_argument("iterations must be positive");
  if (c.chunk_size == 0 || c.chunk_size > c.iterations)
    throw std::out_of_range("\"chunk_size\" must be in [1, iterations]");
  auto job = std::make_unique<sqsg_job>();
  job->iteration_mode = c.iteration_mode;
  job->sublattice_mode = c.sublattice_mode;
  job->num_sites = N;
  job->iterations = c.iterations;
  job->chunk_size = c.chunk_size;
  {  // pool.detach_blocks(0, iterations, worker, iterations / chunk_size) (sqs.h:461-466):
     // BS::thread_pool's near-equal blocks, the first `iterations % blocks` one larger
    const uint64_t nb = std::max<uint64_t>(1, c.iterations / c.chunk_size);
    job->blk_size = c.iterations / nb;
    job->blk_split = (c.iterations % nb) * (job->blk_size + 1);
  }
  job->keep = c.keep;
  job->max_results = c.max_results_per_objective;
  if (c.precision != SQSG_PREC_DOUBLE && c.precision != SQSG_PREC_SINGLE)
    throw std::runtime_error("Unsupported precision");
  job->single = c.precision == SQSG_PREC_SINGLE;
  Lattice lat;
  std::memcpy(lat.m, c.lattice, sizeof(lat.m));
  // apply_composition (structure.h:390-399): sites in ascending order, species by Z
  std::vector<std::vector<uint32_t>> sub_sites(nsub);
  std::vector<uint32_t> owner(N, nsub);
  job->structure_z.assign(c.species, c.species + N);
  for (uint32_t q = 0; q < nsub; ++q) {
    sub_sites[q].assign(c.sites + c.site_offsets[q], c.sites + c.site_offsets[q + 1]);
    if (sub_sites[q].empty()) throw std::out_of_range("You cannot specify an empty slice");
    for (size_t k = 0; k < sub_sites[q].size(); ++k) {
      const uint32_t s = sub_sites[q][k];
      if (s >= N) throw std::out_of_range("site index out of range");
      if (k && s <= sub_sites[q][k - 1]) throw std::invalid_argument("sites must be ascending");
      if (owner[s] != nsub)
        throw std::invalid_argument("The site is contained in more than one sublattice");
      owner[s] = q;
    }
    size_t idx = 0, total = 0;
    for (uint32_t e = c.comp_offsets[q]; e < c.comp_offsets[q + 1]; ++e) total += c.comp_amounts[e];
    if (total != sub_sites[q].size())
      throw std::out_of_range("The total number of distributed atoms does not match the sites");
    for (uint32_t e = c.comp_offsets[q]; e < c.comp_offsets[q + 1]; ++e)
      for (uint32_t a = 0; a < c.comp_amounts[e]; ++a)
        job->structure_z[sub_sites[q][idx++]] = c.comp_species[e];
  }
  const uint32_t nprob = split ? nsub : 1;
  if (nprob > static_cast<uint32_t>(kMaxProblems))
    throw std::out_of_range("split mode supports at most 4 sublattices");
  job->problems.resize(nprob);
  if (!split) {
    // compute_shuffling_bounds (optimization.h:35-72): std::sort by sublattice index
    struct Site {
      uint32_t index, sl;
    };
    std::vector<Site> sites(N);
    for (uint32_t i = 0; i < N; ++i) sites[i] = {i, owner[i]};
    std::sort(sites.begin(), sites.end(),
              [](const Site &a, const Site &b) { return a.sl < b.sl; });
    ProblemInput in;
    in.frac.resize(3 * size_t(N));
    in.z.resize(N);
    in.sort_order.resize(N);
    in.sites.resize(N);
    for (uint32_t k = 0; k < N; ++k) {
      const uint32_t o = sites[k].index;
      in.sort_order[k] = o;
      in.sites[k] = k;
      in.z[k] = job->structure_z[o];
      for (int d = 0; d < 3; ++d) in.frac[3 * k + d] = c.frac_coords[3 * size_t(o) + d];
    }
    uint32_t lower = 0;
    for (uint32_t i = 0; i < N; ++i) {
      if (sites[i].sl != sites[lower].sl) {
        in.bounds.push_back(lower);
        in.bounds.push_back(i);
        lower = i;
        if (sites[i].sl == nsub) break;
      } else if (i == N - 1) {
        in.bounds.push_back(lower);
        in.bounds.push_back(i + 1);
      }
    }
    if (in.bounds.size() / 2 != nsub) throw std::invalid_argument("Incompatible configuration");
    build_problem(job->problems[0], in, lat, c, 0, systematic);
  } else {
    for (uint32_t q = 0; q < nsub; ++q) {
      ProblemInput in;
      const uint32_t n = static_cast<uint32_t>(sub_sites[q].size());
      in.frac.resize(3 * size_t(n));
      in.z.resize(n);
      in.sort_order.resize(n);
      for (uint32_t k = 0; k < n; ++k) {
        const uint32_t o = sub_sites[q][k];
        in.sort_order[k] = k;
        in.z[k] = job->structure_z[o];
        for (int d = 0; d < 3; ++d) in.frac[3 * k + d] = c.frac_coords[3 * size_t(o) + d];
      }
      in.sites = sub_sites[q];
      in.bounds = {0, n};
      build_problem(job->problems[q], in, lat, c, q, false);
    }
  }
  job->stride = 0;
  job->cube_total = 0;
  job->rank_fast = true;
  for (const auto &P : job->problems) {
    job->stride += P.N;
    job->cube_total += size_t(P.S) * P.K * P.K;
  }
  // reported ranks (results.h:163-169) fit 128-bit arithmetic when multinomial * n < 2^128
  auto fits = [](const BigUint &m, size_t n) {
    BigUint t(m);
    t.mul_small(static_cast<uint32_t>(n));
    return t.fits_u128();
  };
  for (const auto &P : job->problems) {
    std::vector<uint8_t> orig(P.N);
    for (uint32_t k = 0; k < P.N; ++k) orig[k] = P.species_packed[k];
    if (!fits(num_permutations(orig.data(), orig.size()), orig.size())) job->rank_fast = false;
  }
  if (systematic) {
    const Problem &P = job->problems[0];
    const BigUint total =
        num_permutations(P.species_packed.data() + P.bounds[0], P.bounds[1] - P.bounds[0]);
    if (!total.fits_u64())
      throw std::out_of_range("The number of permutations to test is " + total.str() +
                              ", more than 2^64 - 1");
    if (static_cast<u128>(c.iterations) > total.to_u128())
      throw std::out_of_range("iterations exceed the number of permutations");
  }
  return job.release();
}

std::string kernel_shape(sqsg_job *job) {
  analyze_job(job);
  switch (resolve_kernel(job, job->kernel)) {
    case SQSG_KERNEL_DENSE_MFMA: {
      const uint32_t TT = dense_tile_bound_of(job), alt = search_alt(job);
      std::string s = "dense<" +
                      (alt ? dense_alt_shape_string(TT, alt) : dense_shape_string(TT, false)) +
                      ">";
      s += " columns ";
      for (size_t q = 0; q < job->problems.size(); ++q)
        s += (q ? "+" : "") + std::to_string(job->problems[q].Kc);
      if (job->dense_passes) s += " passes " + std::to_string(job->dense_passes);
      return s;
    }
    case SQSG_KERNEL_DENSE_PAIRED:
      return "paired<" + dense_shape_string(dense_tile_bound_of(job), true) + ">";
    case SQSG_KERNEL_SPARSE_LANE:
      return "lane<" + std::to_string(job->kmax) + ">";
    default:
      return "wave";
  }
}

bool paired_ok(const sqsg_job *job) {
  return job->dense_ok && job->problems.size() == 1 &&
         job->iteration_mode != SQSG_ITERATION_SYSTEMATIC &&
         dense_paired_available(dense_tile_bound_of(job));
}

int resolve_kernel(const sqsg_job *job, int requested) {
  switch (requested) {
    case SQSG_KERNEL_AUTO:  // exact everywhere; fastest first where it applies
      return job->dense_ok  ? SQSG_KERNEL_DENSE_MFMA
             : job->lane_ok ? SQSG_KERNEL_SPARSE_LANE
                            : SQSG_KERNEL_SPARSE_WAVE;
    case SQSG_KERNEL_SPARSE_WAVE:
      return requested;
    case SQSG_KERNEL_SPARSE_LANE:
      if (!job->lane_ok)
        throw std::invalid_argument("the lane kernel needs at most 8 species and fewer than "
                                    "65536 pairs per shell");
      return requested;
    case SQSG_KERNEL_DENSE_MFMA:
      if (!job->dense_ok) throw std::invalid_argument(kDenseDomain);
      return requested;
    case SQSG_KERNEL_DENSE_PAIRED:
      if (!paired_ok(job))
        throw std::invalid_argument("the paired dense kernel needs a Monte Carlo job the dense "
                                    "kernel handles, of at most 256 sites");
      return requested;
    default:
      throw std::invalid_argument("kernel variant not available for this job");
  }
}

// The wavefront filter bounds records where the reference's insertion of a candidate depends
// on the objectives before it alone: not in systematic runs (a wavefront's lanes walk
// interleaved rank ranges), not where a skip for max_results_per_objective stalls an interact
// block (the stalled candidates' objectives are not the reference's).
bool records_bounded(const sqsg_job *job) {
  return job->iteration_mode != SQSG_ITERATION_SYSTEMATIC &&
         job->keep + 1 <= static_cast<uint64_t>(kMaxLevels) &&
         (job->max_results < 0 || job->sublattice_mode == SQSG_SUBLATTICE_SPLIT);
}

uint32_t first_sink_records(const sqsg_job *job, uint64_t waves, uint64_t launch) {
  // with the filter, a wavefront records about L ln(n) + ties of its n candidates while the
  // threshold is still loose; without it, launches start short and the sink grows on demand
  const uint64_t want = records_bounded(job)
                            ? std::max<uint64_t>(uint64_t(1) << 14, waves * (job->keep + 1) * 16)
                            : uint64_t(1) << 16;
  return static_cast<uint32_t>(std::min<uint64_t>({launch, want, uint64_t(1) << 22}));
}

uint32_t search_alt(const sqsg_job *job) {
  return job->problems.size() == 1 ? job->dense_alt : 0u;
}

RunParams make_run_params(sqsg_job *job, int variant, uint32_t W, uint32_t *blocks,
                          bool search) {
  RunParams p{};
  p.nprob = static_cast<uint32_t>(job->problems.size());
  uint32_t stride_sum = 0, max_bins = 0, max_terms = 0;
  for (uint32_t q = 0; q < p.nprob; ++q) {
    const Problem &P = job->problems[q];
    p.prob[q] = P.dev();
    stride_sum += P.row_stride();
    max_bins = std::max(max_bins, P.S * P.K * P.K);
    max_terms = std::max(max_terms, static_cast<uint32_t>(P.terms.size()));
  }
  p.systematic = job->iteration_mode == SQSG_ITERATION_SYSTEMATIC;
  p.chunk_quirk = (!p.systematic && job->sublattice_mode == SQSG_SUBLATTICE_INTERACT) ? 1u : 0u;
  p.blk_size = job->blk_size;
  p.blk_split = job->blk_split;
  p.L = static_cast<uint32_t>(std::min<uint64_t>(job->keep + 1, kMaxLevels));
  p.wave_filter = records_bounded(job) ? 1u : 0u;
  p.prune = job->prune ? 1u : 0u;
  auto align = [](uint32_t x) { return (x + 15u) & ~15u; };
  uint32_t off = 0;
  if (variant == SQSG_KERNEL_DENSE_PAIRED) {
    // per pair: two row buffers of gen_group candidates (the batch being scored and the
    // next one being generated) and the counts of two shells.  The layout depends on the
    // tile bound and K only, so that the jobs of a persistent launch (grouped by both)
    // share it.
    const Problem &P = job->problems[0];
    const uint32_t TT = dense_tile_bound((P.N + 31) / 32), G = dense_groups(TT);
    const uint32_t g = std::min(32 / P.K, 64 / G);
    p.dense_cw = G * g;
    p.dense_stride = 32 * TT + 16;
    const uint32_t cnt_bytes = align(p.dense_cw * P.K * P.K * 4);
    const uint32_t budget = (160 * 1024 - kDensePairHead) / kDensePairs;
    uint32_t npass = 64 / p.dense_cw;
    while (npass > 1 && 2 * align(npass * p.dense_cw * p.dense_stride) + 2 * cnt_bytes > budget)
      --npass;
    p.gen_group = npass * p.dense_cw;
    p.off_rows[0] = 0;
    p.off_rows[1] = align(p.gen_group * p.dense_stride);
    p.off_hist = 2 * p.off_rows[1];
    p.off_terms = p.off_hist + cnt_bytes;
    p.ws_pair = p.off_terms + cnt_bytes;
    p.wave_lds = kDensePairHead + kDensePairs * p.ws_pair;
    if (p.wave_lds > 160 * 1024) throw std::out_of_range("problem does not fit the LDS budget");
    *blocks = static_cast<uint32_t>(device_cu_count()) *
              dense_paired_blocks_per_cu(TT, p.wave_lds, false);
    return p;
  }
  if (variant == SQSG_KERNEL_DENSE_MFMA) {
    // per problem rows of gen_group candidates (npass passes of dense_cw), then one shell's
    // counts; split jobs: the problems share the tile bound, the row stride and the pass
    // width of the largest count of one-hot columns, and each lane's objectives per problem
    // follow.  The search kernels take a problem's Kc columns per candidate, explicit
    // scoring its K.
    ensure_dense(job);
    for (uint32_t q = 0; q < p.nprob; ++q) p.prob[q] = job->problems[q].dev();
    const uint32_t nprob = p.nprob;
    uint32_t T = 0, Kmax = 0, Cmax = 0;
    for (const auto &P : job->problems) {
      T = std::max(T, (P.N + 31) / 32);
      Kmax = std::max(Kmax, P.K);
      Cmax = std::max(Cmax, search ? P.Kc : P.K);
    }
    const uint32_t TT = dense_tile_bound(T), alt = search ? search_alt(job) : 0u;
    const uint32_t G = alt ? dense_alt_groups(TT, alt) : dense_groups(TT);
    const uint32_t g = std::min(32 / Cmax, 64 / G);
    p.dense_cw = G * g;
    p.dense_stride = 32 * T + 16;
    // one shell's counts of a pass and 32 spare words; then the tie filter levels
    const uint32_t cnt_bytes = align(p.dense_cw * (Kmax * Kmax | 1u) * 4 + 32 * 4);
    const uint32_t lf_bytes = align(static_cast<uint32_t>(sizeof(LevelFilter)));
    const uint32_t obj_bytes = nprob > 1 ? align(nprob * 64 * 8) : 0;
    auto rows_bytes = [&](uint32_t npass) {
      const uint32_t r = npass * p.dense_cw * p.dense_stride;
      return nprob == 1 ? r : nprob * align(r);
    };
    uint32_t npass = 64 / p.dense_cw;
    // at most 20 KiB per wavefront: eight resident wavefronts per CU (two per SIMD)
    while (npass > 1 && rows_bytes(npass) + cnt_bytes + obj_bytes + lf_bytes > 20 * 1024)
      --npass;
    if (search && job->dense_passes) npass = std::min(npass, job->dense_passes);
    p.gen_group = npass * p.dense_cw;
    const uint32_t rb = align(p.gen_group * p.dense_stride);
    for (uint32_t q = 0; q < nprob; ++q) p.off_rows[q] = q * rb;
    p.off_hist = nprob * rb;
    p.off_obj = p.off_hist + cnt_bytes;
    p.off_levels = p.off_obj + obj_bytes;
    p.wave_lds = p.off_levels + lf_bytes;
    if (p.wave_lds > 160 * 1024) throw std::out_of_range("problem does not fit the LDS budget");
    p.prune = job->prune && nprob == 1 ? 1u : 0u;  // the early exit bounds one problem's terms
    // one wavefront per block, as many as registers and LDS keep resident
    *blocks = static_cast<uint32_t>(device_cu_count()) *
              (alt ? dense_alt_waves_per_cu(TT, alt, p.wave_lds, false)
                   : dense_waves_per_cu(TT, p.wave_lds, false, nprob > 1));
    return p;
  }
  if (variant == SQSG_KERNEL_SPARSE_LANE) {
    p.gen_group = 64;
    for (uint32_t q = 0; q < p.nprob; ++q) {
      p.off_rows[q] = off;
      off = align(off + 64 * (job->problems[q].N + 1));
    }
    p.off_hist = off;
    off = align(off + job->kmax * ((job->kmax + 1) / 2) * 64 * 4);
  } else {
    uint32_t G = 64;
    while (G > 8 && static_cast<size_t>(G) * stride_sum > 24576) G >>= 1;
    p.gen_group = G;
    for (uint32_t q = 0; q < p.nprob; ++q) {
      p.off_rows[q] = off;
      off = align(off + G * job->problems[q].row_stride());
    }
    p.off_hist = off;
    off = align(off + 4 * max_bins);
    p.off_terms = off;
    off = align(off + 8 * max_terms);
    p.off_obj = off;
    off = align(off + 8 * 64);
  }
  p.wave_lds = off;
  const size_t block_lds = static_cast<size_t>(W) * p.wave_lds;
  if (block_lds > 160 * 1024) throw std::out_of_range("problem does not fit the LDS budget");
  uint32_t per_cu = static_cast<uint32_t>((160 * 1024) / block_lds);
  per_cu = std::max(1u, std::min(per_cu, 16u / W));
  *blocks = static_cast<uint32_t>(device_cu_count()) * per_cu;
  return p;
}

hipError_t launch_variant(const RunParams &p, const sqsg_job *job, int variant, uint32_t blocks,
                          uint32_t W, hipStream_t s) {
  if (variant == SQSG_KERNEL_SPARSE_LANE) return launch_run_lane(p, job->kmax, blocks, W, s);
  if (variant == SQSG_KERNEL_DENSE_MFMA) {
    if (const uint32_t alt = search_alt(job))
      return launch_run_dense_alt(p, dense_tile_bound_of(job), alt, blocks, s);
    return launch_run_dense(p, dense_tile_bound_of(job), blocks, s, job->problems.size() > 1);
  }
  if (variant == SQSG_KERNEL_DENSE_PAIRED)
    return launch_run_dense_paired(p, dense_tile_bound_of(job), blocks, s);
  return launch_run_wave(p, blocks, W, s);
}

void RecordSink::ensure(uint32_t n, uint32_t stride, uint32_t nprob) {
  // buffers are kept when large enough (a pooled sink may come from a job of another shape)
  if (cand_buf.size() < size_t(n) * sizeof(uint64_t)) cand_buf.alloc(size_t(n) * sizeof(uint64_t));
  if (obj_buf.size() < size_t(n) * sizeof(double)) obj_buf.alloc(size_t(n) * sizeof(double));
  if (sub_buf.size() < size_t(n) * nprob * sizeof(double))
    sub_buf.alloc(size_t(n) * nprob * sizeof(double));
  if (species_buf.size() < size_t(n) * stride) species_buf.alloc(size_t(n) * stride);
  cand = cand_buf.as<uint64_t>();
  obj = obj_buf.as<double>();
  sub = sub_buf.as<double>();
  species = species_buf.as<uint8_t>();
  cap = n;
}

uint32_t RecordSink::place(void *base, size_t bytes, uint32_t stride, uint32_t nprob) {
  const size_t per = sizeof(uint64_t) + sizeof(double) * (1 + nprob) + stride;
  const size_t n = bytes > 1024 ? (bytes - 1024) / per : 0;  // alignment slack of 4 arrays
  cap = static_cast<uint32_t>(std::min<size_t>(n, UINT32_MAX));
  auto at = [](uintptr_t x) { return (x + 255) & ~uintptr_t(255); };
  uintptr_t o = at(reinterpret_cast<uintptr_t>(base));
  cand = reinterpret_cast<uint64_t *>(o);
  o = at(o + size_t(cap) * sizeof(uint64_t));
  obj = reinterpret_cast<double *>(o);
  o = at(o + size_t(cap) * sizeof(double));
  sub = reinterpret_cast<double *>(o);
  o = at(o + size_t(cap) * nprob * sizeof(double));
  species = reinterpret_cast<uint8_t *>(o);
  return cap;
}

void RecordSink::bind(RunParams &p) const {
  p.rec_cap = cap;
  p.rec_cand = cand;
  p.rec_obj = obj;
  p.rec_sub = sub;
  p.rec_species = species;
}

std::shared_ptr<RecordBlock> RecordSink::fetch(const sqsg_job *job, uint32_t count,
                                               uint32_t nprob, hipStream_t s) const {
  const uint32_t stride = job->stride;
  auto blk = std::make_shared<RecordBlock>();
  blk->cand.resize(count);
  blk->obj.resize(count);
  blk->sub.resize(size_t(count) * nprob);
  blk->species.resize(size_t(count) * stride);
  if (count) {
    SQSG_HIP(hipMemcpyAsync(blk->cand.data(), cand, count * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, s));
    SQSG_HIP(hipMemcpyAsync(blk->obj.data(), obj, count * sizeof(double), hipMemcpyDeviceToHost,
                            s));
    SQSG_HIP(hipMemcpyAsync(blk->sub.data(), sub, blk->sub.size() * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    SQSG_HIP(hipMemcpyAsync(blk->species.data(), species, blk->species.size(),
                            hipMemcpyDeviceToHost, s));
    SQSG_HIP(hipStreamSynchronize(s));
  }
  return blk;
}

void RecordSink::drain(sqsg_job *job, uint32_t count, uint32_t nprob,
                       const std::vector<std::pair<uint64_t, uint64_t>> &ranges,
                       hipStream_t s) {
  merge_block(job, fetch(job, count, nprob, s), ranges);
}

void merge_block(sqsg_job *job, std::shared_ptr<const RecordBlock> blk,
                 const std::vector<std::pair<uint64_t, uint64_t>> &ranges) {
  // the threshold only decreases as candidates are merged, so a record above it now is
  // above it at its turn: it is dropped before the sort
  const double search = job->coll.search;
  const uint32_t count = static_cast<uint32_t>(blk->cand.size());
  for (const auto &[begin, end] : ranges) {
    HostRecords r;
    r.begin = begin;
    r.end = end;
    r.block = blk;
    for (uint32_t k = 0; k < count; ++k)
      if (blk->cand[k] >= begin && blk->cand[k] < end && blk->obj[k] <= search)
        r.order.emplace_back(blk->cand[k], k);
    std::sort(r.order.begin(), r.order.end());
    merge_records(job, std::move(r));
  }
}

std::shared_ptr<RecordBlock> stage_block(const sqsg_job *job, const uint8_t *stage,
                                         uint32_t cap, uint32_t count, uint32_t nprob) {
  const uint32_t stride = job->stride;
  auto blk = std::make_shared<RecordBlock>();
  const uint64_t *cand = reinterpret_cast<const uint64_t *>(stage + 16);
  const double *obj = reinterpret_cast<const double *>(cand + cap);
  const double *sub = obj + cap;
  const uint8_t *sp = reinterpret_cast<const uint8_t *>(sub + size_t(cap) * nprob);
  blk->cand.assign(cand, cand + count);
  blk->obj.assign(obj, obj + count);
  blk->sub.assign(sub, sub + size_t(count) * nprob);
  blk->species.assign(sp, sp + size_t(count) * stride);
  return blk;
}

uint64_t default_step_size(const sqsg_job *job) {
  uint64_t work = 0;
  for (const auto &P : job->problems) work += P.pi.size() + 64;
  const uint64_t launch = (uint64_t(1) << 34) / std::max<uint64_t>(work, 1);
  return std::clamp<uint64_t>(launch, uint64_t(1) << 16, uint64_t(1) << 24);
}

uint64_t default_launch_size(const sqsg_job *job) {
  uint64_t work = 0;
  for (const auto &P : job->problems) work += P.pi.size() + 64;
  // about 2^34 pair visits per launch for the pair-list kernels; the dense kernel covers a
  // pair visit about ten times faster and gets 2^36 (4M candidates at 256 sites, ~5 ms), as
  // each launch ends with its slowest wavefront's last segment and a host merge
  const int v = job->analyzed ? resolve_kernel(job, job->kernel) : SQSG_KERNEL_SPARSE_WAVE;
  const bool dense = v == SQSG_KERNEL_DENSE_MFMA || v == SQSG_KERNEL_DENSE_PAIRED;
  const uint64_t launch = (uint64_t(1) << (dense ? 36 : 34)) / std::max<uint64_t>(work, 1);
  return std::clamp<uint64_t>(launch, uint64_t(1) << 16, uint64_t(1) << 24);
}

void begin_run(sqsg_job *job) {
  ensure_device(job);
  job->cancel.store(false);
  reset_results(job);
  job->stats = sqsg_statistics{};
  job->stats.best_objective = std::numeric_limits<double>::infinity();
  job->kernel_ms = 0;
  job->launches = 0;
  job->scored = 0;
}

void run_job(sqsg_job *job, sqsg_progress_fn cb, void *user, const FixedScratch *fixed) {
  using clock = std::chrono::steady_clock;
  const auto t_start = clock::now();
  analyze_job(job);  // kernel applicability (host data), before the variant is resolved
  const int variant = resolve_kernel(job, job->kernel);
  const bool dense = variant == SQSG_KERNEL_DENSE_MFMA || variant == SQSG_KERNEL_DENSE_PAIRED;
  if (fixed && !job->on_device)
    throw std::invalid_argument("the job's inputs are not resident on the device yet (a first "
                                "sqsg_run of the job uploads them)");
  if (fixed && dense && !job->problems[0].d_dtiles.size())
    throw std::invalid_argument("the job's dense tile stream is not resident on the device yet "
                                "(a first sqsg_run of the job with this kernel uploads it)");
  begin_run(job);
  if (dense) ensure_dense(job);
  const hipStream_t stream = job_stream(job);
  const uint32_t W = variant == SQSG_KERNEL_SPARSE_WAVE ? 4 : 1;
  uint32_t blocks = 0;
  RunParams p = make_run_params(job, variant, W, &blocks);
  const uint64_t launch = job->launch_size    ? job->launch_size
                          : job->tuned_launch ? job->tuned_launch
                                              : default_launch_size(job);
  const uint32_t stride = job->stride;
  p.rec_stride = stride;
  RecordSink sink;
  const bool ramp_up = !records_bounded(job);
  DevBuf d_ctl;
  uint8_t *ctl = nullptr;  // record count (+8), scored count (+16), segment counter (+24)
  if (fixed) {
    if (sink.place(static_cast<uint8_t *>(fixed->base) + 256, fixed->bytes - 256, stride,
                   p.nprob) == 0)
      throw std::out_of_range("a record does not fit the scratch buffer");
    ctl = static_cast<uint8_t *>(fixed->base);
  } else {
    sink.ensure(first_sink_records(job, uint64_t(blocks) * W, launch), stride, p.nprob);
    d_ctl.alloc(256);
    ctl = d_ctl.as<uint8_t>();
  }
  p.thr_key = search_key(job);
  p.rec_count = reinterpret_cast<unsigned int *>(ctl + 8);
  p.scored = reinterpret_cast<unsigned long long *>(ctl + 16);
  p.next_seg = reinterpret_cast<unsigned int *>(ctl + 24);
  p.stagger_from = blocks / 2;
  const auto t_loop = clock::now();
  job->stats.timing_chunk_setup_ns =
      std::chrono::duration_cast<std::chrono::nanoseconds>(t_loop - t_start).count();
  uint64_t begin = 0;
  uint32_t ramp = 0;
  uint64_t limit = launch;  // fixed scratch: launches shrink while their records overflow it
  while (begin < job->iterations) {
    if (job->cancel.load()) break;
    // without the wavefront filter the device records every candidate at or below the
    // threshold as it stood before the launch, which is loose at first: such launches grow
    // from 4096 candidates by 4x up to the launch size
    const uint64_t size = std::min(
        limit, ramp_up ? std::min(launch, uint64_t(4096) << std::min(2 * ramp, 40u)) : launch);
    const uint64_t end = std::min(job->iterations, begin + size);
    if (p.systematic) {
      const uint64_t lanes = uint64_t(blocks) * W * p.gen_group;
      p.rounds = static_cast<uint32_t>(std::clamp<uint64_t>(size / (lanes * 2), 1, 1024));
    } else {
      p.rounds = 1;
    }
    p.cand_begin = begin;
    p.cand_end = end;
    sink.bind(p);
    SQSG_HIP(hipMemsetAsync(ctl + 8, 0, 24, stream));  // record, scored, segment counts
    SQSG_HIP(hipEventRecord(job->ev0, stream));
    SQSG_HIP(launch_variant(p, job, variant, blocks, W, stream));
    SQSG_HIP(hipEventRecord(job->ev1, stream));
    unsigned int count = 0;
    unsigned long long scored = 0;
    SQSG_HIP(hipMemcpyAsync(&count, p.rec_count, sizeof(count), hipMemcpyDeviceToHost,
                            stream));
    SQSG_HIP(hipMemcpyAsync(&scored, p.scored, sizeof(scored), hipMemcpyDeviceToHost,
                            stream));
    SQSG_HIP(hipStreamSynchronize(stream));
    float ms = 0;
    SQSG_HIP(hipEventElapsedTime(&ms, job->ev0, job->ev1));
    job->kernel_ms += ms;
    job->launches++;
    if (count > sink.cap) {  // sink overflow: run the same range again, wider or shorter
      if (!fixed) {
        sink.ensure(count + count / 4, stride, p.nprob);
      } else if (end - begin > 1) {
        limit = (end - begin) / 2;
      } else {
        throw std::out_of_range("the records of one candidate exceed the scratch buffer");
      }
      continue;
    }
    if (scored != end - begin) throw DeviceError("device scored an unexpected candidate count");
    job->scored += scored;
    sink.drain(job, count, p.nprob, {{begin, end}}, stream);
    p.thr_key = search_key(job);
    job->stats.finished = job->scored;
    begin = end;
    ++ramp;
    if (limit < launch) limit = std::min(launch, 2 * limit);
    if (cb) {
      const auto now = clock::now();
      job->stats.timing_loop_ns =
          std::chrono::duration_cast<std::chrono::nanoseconds>(now - t_loop).count();
      if (cb(&job->stats, user)) job->cancel.store(true);
    }
  }
  finalize_results(job, fixed == nullptr);
  const auto t_end = clock::now();
  job->stats.timing_loop_ns =
      std::chrono::duration_cast<std::chrono::nanoseconds>(t_end - t_loop).count();
  job->stats.timing_total_ns =
      std::chrono::duration_cast<std::chrono::nanoseconds>(t_end - t_start).count();
}

void score_batch(sqsg_job *job, uint32_t problem, const uint8_t *d_species, uint32_t B,
                 int32_t *d_bonds, double *d_obj, double *d_sro, int kernel, hipStream_t s) {
  if (problem >= job->problems.size()) throw std::out_of_range("problem index out of range");
  ensure_device(job);
  int variant = kernel == SQSG_KERNEL_AUTO ? SQSG_KERNEL_SPARSE_WAVE
                                           : resolve_kernel(job, kernel);
  // explicit scoring has no generation to overlap: the paired form scores as the dense one
  if (variant == SQSG_KERNEL_DENSE_PAIRED) variant = SQSG_KERNEL_DENSE_MFMA;
  const Problem &P = job->problems[problem];
  ScoreParams sp{};
  sp.prob = P.dev();
  sp.species = d_species;
  sp.B = B;
  sp.bonds = d_bonds;
  sp.objective = d_obj;
  sp.sro = d_sro;
  auto align = [](uint32_t x) { return (x + 15u) & ~15u; };
  if (variant == SQSG_KERNEL_SPARSE_LANE) {
    sp.off_row = 0;
    sp.off_hist = align(64 * (P.N + 1));
    sp.wave_lds = align(sp.off_hist + job->kmax * ((job->kmax + 1) / 2) * 256);
    if (sp.wave_lds > 160 * 1024) throw std::out_of_range("problem exceeds LDS");
    SQSG_HIP(launch_score_lane(sp, job->kmax, s ? s : job_stream(job)));
    return;
  }
  if (variant == SQSG_KERNEL_DENSE_MFMA) {
    ensure_dense(job);
    uint32_t blocks = 0;
    const RunParams rp = make_run_params(job, SQSG_KERNEL_DENSE_MFMA, 1, &blocks, false);
    sp.prob = P.dev();
    sp.dense_cw = rp.dense_cw;
    sp.dense_groups = dense_groups(dense_tile_bound_of(job));
    const uint64_t want = uint64_t(8) * static_cast<uint64_t>(device_cu_count());
    const uint32_t TT = dense_tile_bound_of(job);
    if (dense_shell_available(TT) && (uint64_t(B) + rp.dense_cw - 1) / rp.dense_cw < want) {
      // a batch too small for a pass of whole candidates per wavefront: one wavefront per
      // (pass, shell), the term values in scratch, then their sums
      sp.dense_stride = rp.dense_stride;
      // a wavefront's candidate rows, later its counts
      sp.shell_rows = align(std::max(rp.dense_cw * rp.dense_stride, rp.dense_cw * P.K * P.K * 4));
      sp.off_row = align(dense_shell_tile_lds(TT) +
                         P.K * (P.K + 1) / 2 * static_cast<uint32_t>(sizeof(TermParam)));
      sp.wave_lds = sp.off_row + dense_shell_waves() * sp.shell_rows;
      if (sp.wave_lds > 160 * 1024) throw std::out_of_range("problem exceeds LDS");
      const size_t bytes = size_t(B) * P.terms.size() * sizeof(double);
      if (job->score_scratch.size() < bytes) job->score_scratch.alloc(bytes);
      SQSG_HIP(launch_score_dense_shells(sp, TT, job->score_scratch.as<double>(),
                                         s ? s : job_stream(job)));
      return;
    }
    // candidates per wavefront: whole passes, fewer of them when the batch would otherwise
    // leave most SIMDs idle (aim: two wavefronts per SIMD); a batch that does not fill them
    // with one pass per wavefront runs one tile group per wavefront (twice the wavefronts,
    // each with half the MFMA chain)
    uint32_t gen = rp.gen_group;
    while (gen > rp.dense_cw && (uint64_t(B) + gen - 1) / gen < want) gen -= rp.dense_cw;
    if (gen == rp.dense_cw && sp.dense_groups > 1 && (uint64_t(B) + gen - 1) / gen < want) {
      sp.dense_cw = rp.dense_cw / sp.dense_groups;
      sp.dense_groups = 1;
      gen = sp.dense_cw;
    }
    sp.dense_gen = gen;
    sp.dense_stride = rp.dense_stride;
    sp.off_row = rp.off_rows[0];
    sp.off_hist = rp.off_hist;
    sp.wave_lds = rp.wave_lds;
    SQSG_HIP(launch_score_dense(sp, dense_tile_bound_of(job), s ? s : job_stream(job)));
    return;
  }
  if (variant != SQSG_KERNEL_SPARSE_WAVE)
    throw std::invalid_argument("kernel variant not available for explicit scoring");
  sp.off_row = 0;
  sp.off_hist = align(P.row_stride());
  sp.off_terms = align(sp.off_hist + 4 * P.S * P.K * P.K);
  sp.wave_lds = align(sp.off_terms + 8 * static_cast<uint32_t>(P.terms.size()));
  if (size_t(4) * sp.wave_lds > 160 * 1024) throw std::out_of_range("problem exceeds LDS");
  SQSG_HIP(launch_score_explicit(sp, s ? s : job_stream(job)));
}

}  // namespace sqsg

sqsg_job::~sqsg_job() {
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}
