// Plain-old-data parameter blocks shared by the host driver and the HIP kernels.
#pragma once

#include <cstddef>
#include <cstdint>

namespace sqsg {

constexpr int kMaxProblems = 4;   // split-mode sublattices handled in one candidate
constexpr int kMaxBounds = 8;     // shuffle windows per problem (interact sublattices)
constexpr int kMaxLevels = 16;    // levels of the wavefront filter (keep + 1 at most)
constexpr int kMaxLocalSpecies = 16;
constexpr uint32_t kPairIndexBits = 12;  // packed pair: i | j << 12 | shell << 24
constexpr uint32_t kMaxSites = 1u << kPairIndexBits;

// One objective term (optimization.h:85-100) in reference order: shell s, then xi, then
// eta = xi (the diagonal) and eta > xi; parameters of (s, xi, eta), pw scaled by the shell
// weight.
struct TermParam {
  double pref, pw, tgt;
  uint32_t xi, eta;
};

// One sublattice problem as the kernels see it (optimization_config_data, device side).
struct DevProblem {
  const uint32_t *pairs;      // P packed pairs, i < j in sorted-site order
  const uint8_t *species0;    // N packed species (sorted configuration), padded to 4 bytes
  const double *pref;         // [S][K][K]
  const double *pw;           // [S][K][K] scaled by the shell weights
  const double *tgt;          // [S][K][K]
  const uint32_t *terms;      // T objective terms in reference order: s | xi << 8 | eta << 16
  const TermParam *tparams;   // the same T terms with their parameters
  // dense contraction: 32 x 32 int8 tiles of the upper shell adjacency in the A-operand
  // layout of v_mfma_i32_32x32x32_i8, every tile I <= J < dT of every shell, in (shell,
  // row tile, column tile) order
  const uint32_t *dtiles;     // 256 words per tile
  const uint8_t *species0D;   // sorted configuration, 32 * dT bytes, padding sites 0xff
  // one-hot columns per candidate of the dense search kernels: K, or K - 1 when the bonds of
  // the last species follow from the others' (every shuffle window's sites have equal
  // neighbour counts in each shell); then dsum[s][a] is the sum of the shell-s neighbour
  // counts of the sites holding species a, the same for every candidate
  const uint32_t *dsum;
  uint32_t Kc;
  uint32_t dT;                // 32-site tiles
  uint32_t dtiles_n;
  // lane-per-candidate layout: per shell s, groups [i | count << 16] followed by `count`
  // upper neighbours j of site i; lane_off[s] .. lane_off[s + 1] delimit shell s
  const uint32_t *lane_stream;
  const uint32_t *lane_off;
  const uint8_t *species0F;   // species0 * F (field shift of each species), 4-byte padded
  uint32_t F;                 // bits per packed neighbour-count field (K * F <= 32)
  uint32_t N, K, S, P, T;
  uint32_t nbounds;
  uint32_t bounds[2 * kMaxBounds];
  uint32_t draws;             // RNG draws per candidate: sum over windows of (w - 1)
  uint32_t single;            // configuration<float>: objective terms and sums in float
  uint32_t row_stride;        // bytes of one LDS candidate row (N rounded up, bank-padded)
  uint64_t seed;
  // systematic enumeration of window bounds[0..1]
  uint32_t sys_nspecies;
  uint32_t sys_hist[kMaxLocalSpecies];
  uint8_t sys_rev[kMaxLocalSpecies];
};

// Wave-local state of the tie-keeping filter: the L smallest distinct objectives a
// wavefront has seen for its current job.  A candidate above the L-th of them has L
// smaller distinct objectives before it in the job's candidate order, so the reference
// would not insert it (sqs.h:421 against results.h:78-82).  The dense kernels keep it in
// LDS, one copy per wavefront.
struct LevelFilter {
  double lvl[kMaxLevels];
  uint32_t n;
};

struct RunParams {
  DevProblem prob[kMaxProblems];
  uint32_t nprob;
  uint32_t systematic;
  uint32_t chunk_quirk;        // interact MC: the first candidate of every block is unshuffled
  uint32_t rounds;             // systematic: consecutive ranks walked per lane
  // blocks of the reference's thread pool split (near-equal, the first ones one larger):
  // candidates below blk_split form blocks of blk_size + 1, the rest blocks of blk_size
  uint64_t blk_size, blk_split;
  uint64_t cand_begin, cand_end;
  uint32_t L;                  // distinct objective levels of the wavefront filter (keep + 1)
  // the wavefront filter bounds the records: set where a wavefront sees its candidates in
  // ascending order and the reference's insertion rule depends on objectives alone (Monte
  // Carlo runs other than interact runs with max_results_per_objective; keep + 1 <=
  // kMaxLevels)
  uint32_t wave_filter;
  uint32_t gen_group;          // candidates generated per wave and round (<= 64)
  uint32_t dense_cw;           // dense: candidates per MFMA pass (G groups of 32 / K)
  uint32_t dense_stride;       // dense: bytes of one LDS candidate row
  // LDS layout of one wave (bytes)
  uint32_t wave_lds;
  uint32_t off_rows[kMaxProblems];
  uint32_t off_hist, off_terms, off_obj;  // dense: off_hist holds one shell's counts
  uint32_t off_levels;         // dense: the wavefront's tie filter levels (LevelFilter)
  uint32_t ws_pair;            // dense paired form: LDS bytes of one pair's region
  uint32_t prune;              // dense: stop a pass once no candidate can reach the kept levels
  // the host's search threshold: order key of the keep-th distinct inserted objective as of
  // the candidates merged before this launch was issued
  unsigned long long thr_key;
  // record sink and counters of the launch
  unsigned int *rec_count;
  uint64_t *rec_cand;
  double *rec_obj;
  double *rec_sub;             // nprob objectives per record
  uint8_t *rec_species;        // rec_stride bytes per record
  uint32_t rec_cap;
  uint32_t rec_stride;
  unsigned long long *scored;
  unsigned int *next_seg;      // dense single-job kernel: segment counter (zero at launch)
  uint32_t stagger_from;       // dense search kernels: blocks from this index start later
  // persistent form: host memory (pinned, device-visible) that receives the record and
  // scored counts and the first stage_cap records when the step ends, or null
  uint8_t *stage;
  uint32_t stage_cap;
};

// Layout of a job's stage: {record count u32, pad u32, scored u64} then `cap` candidates
// (u64), objectives (f64), nprob sub-objectives each (f64) and species (rec_stride bytes
// each).
inline size_t stage_bytes(uint32_t cap, uint32_t nprob, uint32_t rec_stride) {
  return 16 + size_t(cap) * (16 + 8 * nprob + rec_stride);
}

// One job's candidate range of a persistent step; wavefronts claim it `chunk` candidates
// (one segment) at a time through the item's claim counter.
struct WorkItem {
  uint32_t job;
  uint32_t chunk;
  uint64_t begin, end;
};

struct BatchParams {
  const RunParams *jobs;       // device table, one per job
  const WorkItem *items;       // device table
  const uint32_t *num_items;   // device scalars {items, jobs}: a captured graph stays valid
  const uint32_t *num_jobs;    // while the tables change between replays
  uint32_t wave_lds;           // LDS of one wavefront: the largest job's layout
  unsigned int *claims;        // per-item chunk counters (zeroed by the reset kernel)
  uint32_t max_items;          // rows of the item table
  uint32_t stagger_from;       // dense search kernels: blocks from this index start later
};

struct ScoreParams {
  DevProblem prob;
  const uint8_t *species;      // B x N
  uint32_t B;
  int32_t *bonds;              // B x S x K x K or null
  double *objective;           // B
  double *sro;                 // B x S x K x K or null
  uint32_t wave_lds, off_row, off_hist, off_terms;
  uint32_t dense_cw, dense_gen, dense_stride;
  uint32_t dense_groups;       // tile groups per wavefront of the dense kernel: 1 or the shape's
  uint32_t shell_rows;         // shell-split kernel: LDS bytes of one wavefront's candidate rows
};

}  // namespace sqsg
