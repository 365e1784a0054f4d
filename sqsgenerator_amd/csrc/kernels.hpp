// Host-callable launchers of the HIP kernels (kernels.hip, dense.hip, batch.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>

#include "device_params.hpp"

namespace sqsg {

// explicit candidates -> ordered int32 bonds, FP64 objective, optional SRO cube
hipError_t launch_score_explicit(const ScoreParams &p, hipStream_t stream);

// generation (MC shuffle / systematic unrank+step) + scoring + tie-keeping threshold,
// one wavefront per candidate (the plain single-job kernel)
hipError_t launch_run_wave(const RunParams &p, uint32_t blocks, uint32_t waves_per_block,
                           hipStream_t stream);

// one lane per candidate; kmax = largest species count of the job's problems (<= 8)
hipError_t launch_run_lane(const RunParams &p, uint32_t kmax, uint32_t blocks,
                           uint32_t waves_per_block, hipStream_t stream);
hipError_t launch_score_lane(const ScoreParams &p, uint32_t kmax, hipStream_t stream);
// persistent multi-job form over a device table of jobs and work items
hipError_t launch_batch_lane(const BatchParams &bp, uint32_t kmax, uint32_t blocks,
                             uint32_t waves_per_block, hipStream_t stream);
hipError_t launch_batch_reset(const BatchParams &bp, uint32_t max_jobs, hipStream_t stream);
// copies every job's counters and first records into its host stage (RunParams::stage)
hipError_t launch_batch_stage(const BatchParams &bp, uint32_t max_jobs, hipStream_t stream);
// raises the dynamic LDS limit of the persistent kernel (called outside graph capture)
hipError_t prepare_batch_lane(uint32_t kmax, size_t lds);
// bits per packed neighbour-count field for a largest species count kmax
uint32_t lane_field_bits(uint32_t kmax);

// dense contraction on the MFMA cores; tile_bound = dense_tile_bound(32-site tiles)
uint32_t dense_tile_bound(uint32_t tiles);
uint32_t dense_groups(uint32_t tile_bound);  // 32-column tile groups per wavefront
// split: the form for split-mode jobs (several sublattice problems per candidate), which
// exists for the tile bounds dense_split_available reports
bool dense_split_available(uint32_t tile_bound);
// resident 1-wave blocks per CU of the dense kernel (registers and LDS), at least 1
uint32_t dense_waves_per_cu(uint32_t tile_bound, size_t wave_lds, bool persistent,
                            bool split = false);
hipError_t launch_run_dense(const RunParams &p, uint32_t tile_bound, uint32_t blocks,
                            hipStream_t stream, bool split = false);
hipError_t launch_batch_dense(const BatchParams &bp, uint32_t tile_bound, uint32_t blocks,
                              hipStream_t stream, bool split = false);
hipError_t prepare_batch_dense(uint32_t tile_bound, size_t lds, bool split = false);
hipError_t launch_score_dense(const ScoreParams &p, uint32_t tile_bound, hipStream_t stream);
// explicit scoring split by shell (batches too small to fill the device; tile bounds with
// dense_shell_available): a grid of dense_shell_columns(B, cw) x S blocks of
// dense_shell_waves() wavefronts, one pass of the shape's tile groups each, then the
// candidates' sums.  `vals` is scratch of B x T doubles (the candidates' term values;
// launches that share it must not overlap).  LDS of a block: the shell's tiles or its chunk
// buffers (dense_shell_tile_lds bytes), its term parameters, then one region of
// ScoreParams::shell_rows per wavefront at ScoreParams::off_row
uint32_t dense_shell_waves();
bool dense_shell_available(uint32_t tile_bound);
uint32_t dense_shell_tile_lds(uint32_t tile_bound);
uint32_t dense_shell_columns(uint64_t B, uint32_t cw);
hipError_t launch_score_dense_shells(const ScoreParams &p, uint32_t tile_bound, double *vals,
                                     hipStream_t stream);
// paired form (MFMA and helper wavefronts, four pairs per block of 512 threads; MC jobs of
// tile bounds whose shape runs two wavefronts per SIMD)
constexpr uint32_t kDensePairs = 4, kDensePairHead = 128;  // pairs per block, LDS header
bool dense_paired_available(uint32_t tile_bound);
// "TT,G,W,D" of the dense shape of a tile bound (paired: "TT,G,D")
std::string dense_shape_string(uint32_t tile_bound, bool paired);
uint32_t dense_paired_blocks_per_cu(uint32_t tile_bound, size_t block_lds, bool persistent);
hipError_t launch_run_dense_paired(const RunParams &p, uint32_t tile_bound, uint32_t blocks,
                                   hipStream_t stream);
hipError_t prepare_batch_dense_paired(uint32_t tile_bound, size_t block_lds);
// alternative search shapes of a tile bound (dense_alt.hip), numbered from 1 (0 is the
// default shape above): the same tile bound with other tile groups or ring depths, for jobs
// of one problem; the autotuner times them against the default
uint32_t dense_alt_shapes(uint32_t tile_bound);
uint32_t dense_alt_groups(uint32_t tile_bound, uint32_t alt);
std::string dense_alt_shape_string(uint32_t tile_bound, uint32_t alt);
uint32_t dense_alt_waves_per_cu(uint32_t tile_bound, uint32_t alt, size_t wave_lds,
                                bool persistent);
hipError_t launch_run_dense_alt(const RunParams &p, uint32_t tile_bound, uint32_t alt,
                                uint32_t blocks, hipStream_t stream);
hipError_t prepare_batch_dense_alt(uint32_t tile_bound, uint32_t alt, size_t lds);
hipError_t launch_batch_dense_alt(const BatchParams &bp, uint32_t tile_bound, uint32_t alt,
                                  uint32_t blocks, hipStream_t stream);
hipError_t launch_batch_dense_paired(const BatchParams &bp, uint32_t tile_bound, uint32_t blocks,
                                     hipStream_t stream);

// order-preserving key of a double (unsigned order == numeric order)
inline unsigned long long objective_key(double d) {
  unsigned long long b;
  __builtin_memcpy(&b, &d, 8);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
inline double key_objective(unsigned long long k) {
  unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  double d;
  __builtin_memcpy(&d, &b, 8);
  return d;
}

}  // namespace sqsg
