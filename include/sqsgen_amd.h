/*
 * sqsgen_amd.h -- C ABI of the MI355X SQS scoring/search core.
 *
 * Drop-in boundary for the optimisation path of dgehringer/sqsgenerator
 * (v0.5.6).  The reference's seam is
 *   sqsgen::run_optimization(variant<configuration<float>, configuration<double>>&&,
 *                            log::level, optional<sqs_callback_t>)   include/sqsgen/sqs.h:625-634
 * whose replaceable unit is the worker loop (sqs.h:348-455) fed by
 * optimization_config_data<T> (include/sqsgen/core/optimization_config.h:33-48).
 * Each entry point below cites the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns an int status (SQSG_OK = 0) and never throws; the
 *     message of the last failure on the calling thread is sqsg_last_error().
 *     Status codes map onto the reference's exception types (see below).
 *   - Cubes are row-major [S][K][K] doubles (the reference stores Eigen ColMajor
 *     tensors; the Python/C++ layers convert).  Species arrays hold atomic
 *     numbers Z where they face the user and packed indices 0..K-1 (species
 *     sorted by Z, core/helpers.h:93-107) where they face the kernels.
 *   - Ranks are arbitrary-precision decimal strings, as rank_t (boost cpp_int,
 *     include/sqsgen/types.h:20) prints them.
 *   - Calls on distinct jobs are thread-safe; a job is not re-entrant.  The
 *     caller owns every input array (they are copied at creation).
 */
#ifndef SQSGEN_AMD_H
#define SQSGEN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SQSG_API __attribute__((visibility("default")))

/* status codes: std::invalid_argument, std::out_of_range, std::domain_error,
 * std::runtime_error (sqs.h:218-221,620; core/results.h:172-192), device and
 * allocation failures, cooperative cancellation (types.h:222-232) */
#define SQSG_OK 0
#define SQSG_ERR_INVALID_ARGUMENT 1
#define SQSG_ERR_OUT_OF_RANGE 2
#define SQSG_ERR_DOMAIN 3
#define SQSG_ERR_RUNTIME 4
#define SQSG_ERR_DEVICE 5
#define SQSG_ERR_NO_MEMORY 6
#define SQSG_ERR_CANCELLED 7

/* enums: include/sqsgen/types.h:62-78 */
#define SQSG_ITERATION_RANDOM 0
#define SQSG_ITERATION_SYSTEMATIC 1
#define SQSG_SUBLATTICE_INTERACT 0
#define SQSG_SUBLATTICE_SPLIT 1
/* precision of the configuration (io/config/combined.h:151-159, configuration<float|double>) */
#define SQSG_PREC_DOUBLE 0
#define SQSG_PREC_SINGLE 1

/* scoring kernel selectors for sqsg_score_batch / sqsg_job_set_kernel.  All variants give
 * identical results (bit-exact bonds, objectives and result packs).  AUTO runs the dense
 * kernel where it applies (one sublattice problem, N <= 1024, K <= 16, at most 127 upper
 * neighbours per site and shell), else the lane kernel (K <= 8), else the wave kernel;
 * explicit scoring (sqsg_score_batch) with AUTO uses the wave kernel. */
#define SQSG_KERNEL_AUTO 0
#define SQSG_KERNEL_SPARSE_WAVE 1   /* one wavefront per candidate, LDS pair histogram */
#define SQSG_KERNEL_SPARSE_LANE 2   /* one lane per candidate, LDS [bin][lane] counters */
#define SQSG_KERNEL_DENSE_MFMA 3    /* int8 one-hot x int8 shell adjacency on MFMA, int32 accumulate */
#define SQSG_KERNEL_DENSE_PAIRED 4  /* the dense contraction with an MFMA and a helper wavefront
                                       paired on each SIMD (Monte Carlo jobs, <= 256 sites) */

typedef struct sqsg_job sqsg_job;
typedef struct sqsg_server sqsg_server;

SQSG_API const char *sqsg_last_error(void);
SQSG_API int sqsg_version(void);
/* Initialises the HIP device (index 0) and reports its name and CU count. */
SQSG_API int sqsg_device_info(char *name, size_t cap, int *compute_units);
SQSG_API int sqsg_device_synchronize(void);

/* ---- host geometry (include/sqsgen/core/structure.h) ---------------------------------- */
/* structure.h:34-69, minimum image over 27 cells; lattice rows a, b, c (row-major) */
SQSG_API int sqsg_distance_matrix(const double *lattice, const double *frac, uint32_t n,
                                  double *out);
/* structure.h:195-207 */
SQSG_API int sqsg_distances_naive(const double *dm, uint32_t n, double atol, double rtol,
                                  double *out, uint32_t cap, uint32_t *count);
/* structure.h:209-258 */
SQSG_API int sqsg_distances_histogram(const double *dm, uint32_t n, double bin_width,
                                      double peak_isolation, double *out, uint32_t cap,
                                      uint32_t *count);
/* structure.h:71-100 */
SQSG_API int sqsg_shell_matrix(const double *dm, uint32_t n, const double *radii,
                               uint32_t num_radii, double atol, double rtol, uint32_t *out);
/* structure.h:136-192 (shells ascending); writes [S][K][K], K = distinct species */
SQSG_API int sqsg_prefactors(const uint32_t *shell_matrix, uint32_t n, const uint32_t *shells,
                             uint32_t num_shells, const uint8_t *species_z, double *out);
/* structure.h:335-366 */
SQSG_API int sqsg_supercell(const double *lattice, const double *frac, const uint8_t *species,
                            uint32_t n, uint32_t a, uint32_t b, uint32_t c, double *lattice_out,
                            double *frac_out, uint8_t *species_out);
/* the same geometry in float, for configuration<float> (structure<float>) */
SQSG_API int sqsg_distance_matrix_f32(const float *lattice, const float *frac, uint32_t n,
                                      float *out);
SQSG_API int sqsg_distances_naive_f32(const float *dm, uint32_t n, float atol, float rtol,
                                      float *out, uint32_t cap, uint32_t *count);
SQSG_API int sqsg_distances_histogram_f32(const float *dm, uint32_t n, float bin_width,
                                          float peak_isolation, float *out, uint32_t cap,
                                          uint32_t *count);
SQSG_API int sqsg_shell_matrix_f32(const float *dm, uint32_t n, const float *radii,
                                   uint32_t num_radii, float atol, float rtol, uint32_t *out);
SQSG_API int sqsg_prefactors_f32(const uint32_t *shell_matrix, uint32_t n, const uint32_t *shells,
                                 uint32_t num_shells, const uint8_t *species_z, float *out);
SQSG_API int sqsg_supercell_f32(const float *lattice, const float *frac, const uint8_t *species,
                                uint32_t n, uint32_t a, uint32_t b, uint32_t c, float *lattice_out,
                                float *frac_out, uint8_t *species_out);

/* ---- multiset permutations (include/sqsgen/core/permutation.h) ------------------------ */
SQSG_API int sqsg_num_permutations(const uint8_t *conf, uint32_t n, char *out, size_t cap);
SQSG_API int sqsg_rank_permutation(const uint8_t *conf, uint32_t n, char *out, size_t cap);
SQSG_API int sqsg_unrank_permutation(uint8_t *conf, uint32_t n, const char *rank);

/* ---- job description: core/config.h:32-48 configuration<T> (T = double or float) ---- */
typedef struct sqsg_config {
  /* structure after the supercell was applied (config.h:15-30) */
  uint32_t num_sites;
  const double *lattice;        /* 9 */
  const double *frac_coords;    /* num_sites x 3 */
  const uint8_t *species;       /* num_sites, Z */
  int32_t sublattice_mode;      /* SQSG_SUBLATTICE_* */
  int32_t iteration_mode;       /* SQSG_ITERATION_* */
  /* composition (types.h:90-93): sublattice q owns sites[site_offsets[q] .. site_offsets[q+1])
   * and distributes comp_amounts[comp_offsets[q] ..] of comp_species (ascending Z) */
  uint32_t num_sublattices;
  const uint32_t *site_offsets;
  const uint32_t *sites;
  const uint32_t *comp_offsets;
  const uint8_t *comp_species;
  const uint32_t *comp_amounts;
  /* per problem p (1 for interact, num_sublattices for split) */
  const uint32_t *radii_offsets;   /* num_problems + 1 */
  const double *radii;
  const uint32_t *shell_offsets;   /* num_problems + 1 */
  const uint32_t *shells;          /* ascending shell indices */
  const double *shell_weights;
  const uint32_t *num_species;     /* K of each problem (cube extent) */
  const double *prefactors;        /* concatenated [S][K][K] */
  const double *pair_weights;      /* concatenated [S][K][K], NOT yet scaled by shell weights */
  const double *target_objective;  /* concatenated [S][K][K] */
  uint64_t iterations;
  uint64_t chunk_size;
  const uint64_t *seeds;           /* per problem */
  const uint8_t *seed_present;     /* per problem; absent -> random_device seed */
  uint64_t keep;
  int64_t max_results_per_objective; /* <0: unlimited */
  /* SQSG_PREC_SINGLE: configuration<float>.  Every floating-point input above must then hold
   * float values (stored as double); geometry, scaled pair weights and objectives are
   * evaluated in float (sqs.h:591-634), results hold float values as double.  Such jobs
   * run on the sparse kernels (wave, lane). */
  int32_t precision;
} sqsg_config;

/* optimization_config<T,Mode>::from_config (optimization_config.h:59-141) plus the
 * device upload.  Validates the configuration like the reference's optimizer ctor. */
SQSG_API int sqsg_job_create(const sqsg_config *cfg, sqsg_job **out);
SQSG_API void sqsg_job_destroy(sqsg_job *job);

typedef struct sqsg_problem_info {
  uint32_t num_sites, num_species, num_shells, num_pairs, num_bounds;
  uint32_t bounds[16];            /* (lo, hi) shuffle windows, sorted-site order */
  uint8_t species_z[64];          /* packed index -> Z */
  uint32_t shell_ids[64];         /* packed shell index -> shell id */
} sqsg_problem_info;
SQSG_API int sqsg_job_num_problems(sqsg_job *job, uint32_t *num_problems);
SQSG_API int sqsg_job_problem_info(sqsg_job *job, uint32_t problem, sqsg_problem_info *info);
/* Read-only inputs of one problem as the kernels see them (optimization_config_data):
 * pairs P x {i, j, shell} (sorted-site order), species_packed N, sort_order N,
 * scaled pair weights / prefactors / target [S][K][K].  Any pointer may be NULL. */
SQSG_API int sqsg_job_problem_arrays(sqsg_job *job, uint32_t problem, uint32_t *pairs,
                                     uint8_t *species_packed, uint32_t *sort_order,
                                     double *prefactors, double *pair_weights, double *target);

/* Scores B explicit candidates of one problem on the GPU: count_bonds + compute_objective
 * (core/optimization.h:74-103).  species: B x N packed indices in sorted-site order (host
 * memory).  bonds: B x S x K x K int32, ordered (a = species[i], b = species[j], i < j)
 * like the reference's cube; objective: B; sro: B x S x K x K or NULL. */
SQSG_API int sqsg_score_batch(sqsg_job *job, uint32_t problem, const uint8_t *species,
                              uint32_t B, int32_t *bonds, double *objective, double *sro,
                              int kernel);
/* Same on device buffers, enqueued on `stream` (hipStream_t, NULL = the job's stream). */
SQSG_API int sqsg_score_batch_device(sqsg_job *job, uint32_t problem, const uint8_t *d_species,
                                     uint32_t B, int32_t *d_bonds, double *d_objective,
                                     double *d_sro, int kernel, void *stream);

/* ---- run: optimizer<T,IMode,SMode>::run (sqs.h:299-580) ------------------------------ */
typedef struct sqsg_statistics {   /* types.h:197-206 */
  uint64_t finished;
  uint64_t working;
  uint64_t best_rank;              /* candidate index of the first best objective */
  double best_objective;
  uint64_t timing_total_ns, timing_chunk_setup_ns, timing_loop_ns, timing_comm_ns;
} sqsg_statistics;
/* progress callback (sqs.h:447-451): called on the calling thread once per device
 * launch; return non-zero to stop (sqs_callback_context::stop, types.h:231) */
typedef int (*sqsg_progress_fn)(const sqsg_statistics *stats, void *user);
SQSG_API int sqsg_job_set_kernel(sqsg_job *job, int kernel);
SQSG_API int sqsg_job_set_launch_size(sqsg_job *job, uint64_t candidates_per_launch);
/* Early exit (no reference counterpart; for callers that want time to solution): with every
 * scaled pair weight >= 0 each objective term is >= 0, so a candidate whose partial objective
 * after some shell already exceeds the job's keep threshold can never be kept.  When enabled,
 * the dense kernels stop a pass of candidates at the first shell after which none of them can
 * be kept.  Result packs are identical; scored counts still include every candidate.  Rejected
 * (SQSG_ERR_INVALID_ARGUMENT) when a pair weight is negative; other kernels score in full. */
SQSG_API int sqsg_job_set_pruning(sqsg_job *job, int enable);
SQSG_API int sqsg_run(sqsg_job *job, sqsg_progress_fn cb, void *user);
/* No counterpart in the reference.  For callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc): sqsg_run without device allocations.  The run's
 * threshold, counters and record sink live in a 16 MiB device array of the library (one per
 * GPU, held until the process ends); a launch whose records exceed it runs again as shorter
 * launches, and a job whose single record exceeds it fails with SQSG_ERR_OUT_OF_RANGE.  The
 * job's inputs must already be resident (a first sqsg_run of the job uploads them), else
 * SQSG_ERR_INVALID_ARGUMENT.  SRO cubes are computed when first read, not by the run.  Calls
 * share the array: within a process they are serialised; the results equal sqsg_run's. */
SQSG_API int sqsg_run_noalloc(sqsg_job *job, sqsg_progress_fn cb, void *user);
/* cooperative cancellation, async-safe (sqs.h:385-396) */
SQSG_API void sqsg_cancel(sqsg_job *job);
SQSG_API int sqsg_job_statistics(sqsg_job *job, sqsg_statistics *stats);
/* device-side timing of the scoring launches of the last run: summed HIP-event
 * milliseconds, number of launches, candidates scored */
SQSG_API int sqsg_job_timing(sqsg_job *job, double *kernel_ms, uint64_t *launches,
                             uint64_t *candidates);

/* ---- results: sqs_result_collection / sqs_result_pack (core/results.h:42-99,336-389) -- */
SQSG_API int sqsg_result_count(sqsg_job *job, uint64_t *num_results, uint64_t *num_objectives);
/* result r in ascending objective order.  species_z: all sites, original order, Z
 * (postprocess_results, sqs.h:94-122); sub_objectives: one per problem; sro: concatenated
 * [S][K][K] per problem; rank: decimal (sqs_result_wrapper::rank, results.h:163-169);
 * candidate: device candidate index (MC: counter-RNG stream index, systematic: rank - 1). */
SQSG_API int sqsg_result_get(sqsg_job *job, uint64_t r, double *objective, uint8_t *species_z,
                             double *sub_objectives, double *sro, char *rank, size_t rank_cap,
                             uint64_t *candidate);
/* internal (sorted-site, packed) species of result r, concatenated per problem */
SQSG_API int sqsg_result_species_packed(sqsg_job *job, uint64_t r, uint8_t *species);
/* objective of each level (ascending) and the levels + 1 offsets of their results */
SQSG_API int sqsg_result_levels(sqsg_job *job, double *objectives, uint64_t *level_start);
/* bulk form of sqsg_result_get for packs with millions of ties: objectives, candidates,
 * original-order species (count x num_sites) and ranks as uint64 (UINT64_MAX where the
 * rank needs more than 64 bits or the job is in split mode); any pointer may be NULL */
SQSG_API int sqsg_result_bulk(sqsg_job *job, uint64_t first, uint64_t count, double *objective,
                              uint64_t *candidate, uint8_t *species_z, uint64_t *rank);
/* SRO cubes of results [first, first + count), concatenated per problem */
SQSG_API int sqsg_result_sro_bulk(sqsg_job *job, uint64_t first, uint64_t count, double *sro);

/* ---- batched persistent form (no reference counterpart: it batches many
 *      run_optimization calls into one device launch) ------------------------------------ */
typedef struct sqsg_batch_report {
  double kernel_ms;        /* HIP-event time of the persistent launches */
  uint64_t candidates;     /* candidates scored */
  uint64_t launches;
  uint32_t work_items;
} sqsg_batch_report;
/* Runs every job to completion in shared persistent launches (graph-captured when
 * `use_graph`); each job's results are then read with sqsg_result_*. */
SQSG_API int sqsg_batch_run(sqsg_job **jobs, uint32_t num_jobs, int use_graph,
                            sqsg_batch_report *report);
/* Autotuner (no reference counterpart): times the job's applicable kernel variants on a probe
 * of its own candidates and, for the dense kernel of a one-problem job, its launch
 * parameters: the alternative launch shapes of its tile bound (tile groups, ring depth), one
 * pass of candidates fewer per wavefront round, then the candidates per sqsg_run launch
 * (unless the caller set a launch size).  Keeps the fastest in the job (variant written to
 * *chosen, its probe rate to *cand_per_s; sqsg_job_kernel_shape and sqsg_job_dense_launch
 * report the rest).  Results are identical for every choice; the probes overwrite the job's
 * results. */
SQSG_API int sqsg_autotune(sqsg_job *job, int *chosen, double *cand_per_s);
/* Dense search launch parameters of a job (no reference counterpart; what the autotuner
 * chooses, settable to reproduce a tuned run): `shape` 0 is the tile bound's default launch
 * shape and 1 .. num_shapes - 1 its alternatives (one-problem jobs), `passes` the candidates
 * per wavefront round in passes of the shape's width (0: as many as its LDS budget holds),
 * `launch` the candidates per sqsg_run launch (0: the default; a launch size set with
 * sqsg_job_set_launch_size takes precedence).  Results are identical for every setting. */
SQSG_API int sqsg_job_set_dense_launch(sqsg_job *job, uint32_t shape, uint32_t passes,
                                       uint64_t launch);
SQSG_API int sqsg_job_dense_launch(sqsg_job *job, uint32_t *num_shapes, uint32_t *shape,
                                   uint32_t *passes, uint64_t *launch);
/* Launch shape of the kernel a run of `job` uses (no reference counterpart: introspection
 * for profiling and reports), e.g. "dense<8,2,2,3>" (tile bound, tile groups, wavefronts
 * per SIMD, ring depth), "paired<8,2,9>", "lane<5>" or "wave"; NUL-terminated, truncated
 * to `len` bytes. */
SQSG_API int sqsg_job_kernel_shape(sqsg_job *job, char *buf, size_t len);

/* ---- serving: continuous batching with device-resident job state --------------------- */
typedef struct sqsg_server_options {
  uint64_t hbm_budget_bytes;     /* admission limit on resident job footprint (0: 80% of the
                                    device memory free when the server is created) */
  uint32_t max_active;           /* admission limit on resident jobs */
  uint64_t slice_candidates;     /* per-job candidates per scheduling step */
} sqsg_server_options;
SQSG_API int sqsg_server_create(const sqsg_server_options *opt, sqsg_server **out);
/* Queues a created job; returns a ticket.  A host thread of the server uploads the job's
 * inputs while steps run.  The server owns nothing: the caller keeps the job and must not
 * touch it until sqsg_server_poll reports it done. */
SQSG_API int sqsg_server_submit(sqsg_server *srv, sqsg_job *job, uint64_t *ticket);
/* One scheduling step: admits queued jobs whose inputs are uploaded, in submission order,
 * skipping those that do not fit the memory left, while the resident-job limit allows; then
 * advances every resident job by one persistent step and retires finished ones; reports the
 * resident and queued job counts after the step.  A job whose preparation failed is marked
 * done and its error returned here. */
SQSG_API int sqsg_server_step(sqsg_server *srv, int *active_jobs, int *queued_jobs);
/* Status of a ticket: done flag and submission, admission and completion times in seconds
 * since server creation (-1 where not reached yet). */
SQSG_API int sqsg_server_poll(sqsg_server *srv, uint64_t ticket, int *done, double *t_submit,
                              double *t_admit, double *t_done);
SQSG_API void sqsg_server_destroy(sqsg_server *srv);

#ifdef __cplusplus
}
#endif
#endif
