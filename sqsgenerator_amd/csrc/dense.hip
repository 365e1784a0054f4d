// Host launchers of the dense MFMA kernels (dense_kernels.hpp) at the default launch shape
// of each tile bound: single-job and persistent search, explicit scoring (whole candidate
// and split by shell) and the paired form; alternative search shapes are in dense_alt.hip.
#include "dense_kernels.hpp"

namespace sqsg {
namespace {

// the paired form exists for the tile bounds up to 8 (256 sites)
template <class Sh> constexpr bool ws_shape() { return Sh::W == 2 && Sh::TT <= 8; }
template <class Sh> constexpr int ws_ring() { return Sh::DS; }

template <class Sh> const void *ws_fn(bool persistent) {
  return persistent ? reinterpret_cast<const void *>(&batch_dense_ws_kernel<Sh::TT, Sh::G, ws_ring<Sh>()>)
                    : reinterpret_cast<const void *>(&run_dense_ws_kernel<Sh::TT, Sh::G, ws_ring<Sh>()>);
}

template <int TT, int G, int DS> hipError_t score_dense_g(const ScoreParams &p, hipStream_t s) {
  const void *fn = reinterpret_cast<const void *>(&score_dense_kernel<TT, G, DS>);
  const hipError_t e = dense_lds(fn, p.wave_lds);
  if (e != hipSuccess) return e;
  const uint32_t blocks = static_cast<uint32_t>((p.B + p.dense_gen - 1) / p.dense_gen);
  hipLaunchKernelGGL((score_dense_kernel<TT, G, DS>), dim3(blocks), dim3(64), p.wave_lds, s,
                     p);
  return hipGetLastError();
}

// the shape's tile groups, or one group per wavefront for batches too small to fill the
// device (the host then passes a one-group pass width)
template <class Sh> hipError_t score_dense_k(const ScoreParams &p, hipStream_t s) {
  if constexpr (Sh::G > 1) {
    if (p.dense_groups == 1) return score_dense_g<Sh::TT, 1, Sh::DS>(p, s);
  }
  if (p.dense_groups != static_cast<uint32_t>(Sh::G)) return hipErrorInvalidValue;
  return score_dense_g<Sh::TT, Sh::G, Sh::DS>(p, s);
}

// the shell-split form: a shell's tiles in LDS at once for tile bounds up to 8 (36 KiB),
// streamed through LDS in chunks beyond
template <class Sh> constexpr bool shell_lds_tiles() { return Sh::TT <= 8; }

template <class Sh> hipError_t score_shell_k(const ScoreParams &p, double *vals, hipStream_t s) {
  constexpr bool LT = shell_lds_tiles<Sh>();
  const void *fn = reinterpret_cast<const void *>(&score_shell_kernel<Sh::TT, Sh::G, Sh::DS, LT>);
  const hipError_t e = dense_lds(fn, p.wave_lds);
  if (e != hipSuccess) return e;
  const dim3 grid(dense_shell_columns(p.B, p.dense_cw), p.prob.S);
  hipLaunchKernelGGL((score_shell_kernel<Sh::TT, Sh::G, Sh::DS, LT>), grid,
                     dim3(64 * kShellWaves), p.wave_lds, s, p, vals);
  const hipError_t r = hipGetLastError();
  if (r != hipSuccess) return r;
  hipLaunchKernelGGL(score_sum_kernel, dim3(static_cast<uint32_t>((p.B + kSumCands - 1) / kSumCands)),
                     dim3(kSumThreads), 0, s, p, vals);
  return hipGetLastError();
}

}  // namespace

uint32_t dense_shell_waves() { return kShellWaves; }

bool dense_shell_available(uint32_t tile_bound) { return dense_groups(tile_bound) > 0; }

uint32_t dense_shell_tile_lds(uint32_t tile_bound) {
  uint32_t b = 0;
  (void)with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    b = (shell_lds_tiles<Sh>() ? Sh::TT * (Sh::TT + 1) / 2
                                : kShellBufs * shell_chunk<Sh::TT, Sh::G>()) * 1024u;
    return hipSuccess;
  });
  return b;
}

uint32_t dense_shell_columns(uint64_t B, uint32_t cw) {
  const uint64_t passes = (B + cw - 1) / cw;
  return static_cast<uint32_t>((passes + kShellWaves - 1) / kShellWaves);
}

hipError_t launch_score_dense_shells(const ScoreParams &p, uint32_t tile_bound, double *vals,
                                     hipStream_t s) {
  if (p.B == 0) return hipSuccess;
  return with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    if (p.dense_groups != static_cast<uint32_t>(Sh::G)) return hipErrorInvalidValue;
    return score_shell_k<Sh>(p, vals, s);
  });
}

uint32_t dense_tile_bound(uint32_t tiles) {
  for (uint32_t tt : {1u, 2u, 4u, 6u, 8u, 12u, 14u, 16u, 20u, 24u, 27u, 32u})
    if (tiles <= tt) return tt;
  return 0;
}
uint32_t dense_groups(uint32_t tile_bound) {
  uint32_t g = 0;
  (void)with_dense_shape(tile_bound, [&](auto sh) {
    g = decltype(sh)::G;
    return hipSuccess;
  });
  return g;
}

uint32_t dense_waves_per_cu(uint32_t tile_bound, size_t wave_lds, bool persistent,
                            bool split) {
  int n = 0;
  const hipError_t e = with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    return with_split<Sh>(split, [&](auto sp) {
      constexpr bool SPLIT = decltype(sp)::value;
      const void *fn = persistent ? batch_dense_fn<Sh, SPLIT>() : run_dense_fn<Sh, SPLIT>();
      hipError_t r = dense_lds(fn, wave_lds);
      if (r != hipSuccess) return r;
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64, wave_lds);
    });
  });
  return e == hipSuccess && n > 0 ? static_cast<uint32_t>(n) : 1u;
}

bool dense_split_available(uint32_t tile_bound) {
  bool ok = false;
  (void)with_dense_shape(tile_bound, [&](auto sh) {
    ok = split_shape<decltype(sh)>();
    return hipSuccess;
  });
  return ok;
}

hipError_t launch_run_dense(const RunParams &p, uint32_t tile_bound, uint32_t blocks,
                            hipStream_t s, bool split) {
  return with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    return run_dense_k<Sh>(p, blocks, s, split);
  });
}

hipError_t prepare_batch_dense(uint32_t tile_bound, size_t lds, bool split) {
  return with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    return with_split<Sh>(split, [&](auto sp) {
      return dense_lds(batch_dense_fn<Sh, decltype(sp)::value>(), lds);
    });
  });
}

hipError_t launch_batch_dense(const BatchParams &bp, uint32_t tile_bound, uint32_t blocks,
                              hipStream_t s, bool split) {
  return with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    return batch_dense_k<Sh>(bp, blocks, s, split);
  });
}

hipError_t launch_score_dense(const ScoreParams &p, uint32_t tile_bound, hipStream_t s) {
  if (p.B == 0) return hipSuccess;
  return with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    return score_dense_k<Sh>(p, s);
  });
}

std::string dense_shape_string(uint32_t tile_bound, bool paired) {
  std::string s;
  (void)with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    s = std::to_string(Sh::TT) + "," + std::to_string(Sh::G) + ",";
    s += paired ? std::to_string(ws_ring<Sh>())
                : std::to_string(Sh::W) + "," + std::to_string(Sh::D);
    return hipSuccess;
  });
  return s;
}

bool dense_paired_available(uint32_t tile_bound) {
  bool ok = false;
  (void)with_dense_shape(tile_bound, [&](auto sh) {
    ok = ws_shape<decltype(sh)>();
    return hipSuccess;
  });
  return ok;
}

uint32_t dense_paired_blocks_per_cu(uint32_t tile_bound, size_t block_lds, bool persistent) {
  int n = 0;
  const hipError_t e = with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    if constexpr (!ws_shape<Sh>()) {
      return hipErrorInvalidValue;
    } else {
      const void *fn = ws_fn<Sh>(persistent);
      hipError_t r = dense_lds(fn, block_lds);
      if (r != hipSuccess) return r;
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64 * 2 * kWsPairs, block_lds);
    }
  });
  return e == hipSuccess && n > 0 ? static_cast<uint32_t>(n) : 1u;
}

hipError_t launch_run_dense_paired(const RunParams &p, uint32_t tile_bound, uint32_t blocks,
                                   hipStream_t s) {
  return with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    if constexpr (!ws_shape<Sh>()) {
      return hipErrorInvalidValue;
    } else {
      const hipError_t e = dense_lds(ws_fn<Sh>(false), p.wave_lds);
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL((run_dense_ws_kernel<Sh::TT, Sh::G, ws_ring<Sh>()>), dim3(blocks),
                         dim3(64 * 2 * kWsPairs), p.wave_lds, s, p);
      return hipGetLastError();
    }
  });
}

hipError_t prepare_batch_dense_paired(uint32_t tile_bound, size_t block_lds) {
  return with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    if constexpr (!ws_shape<Sh>()) {
      return hipErrorInvalidValue;
    } else {
      return dense_lds(ws_fn<Sh>(true), block_lds);
    }
  });
}

hipError_t launch_batch_dense_paired(const BatchParams &bp, uint32_t tile_bound, uint32_t blocks,
                                     hipStream_t s) {
  return with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    if constexpr (!ws_shape<Sh>()) {
      return hipErrorInvalidValue;
    } else {
      hipLaunchKernelGGL((batch_dense_ws_kernel<Sh::TT, Sh::G, ws_ring<Sh>()>), dim3(blocks),
                         dim3(64 * 2 * kWsPairs), bp.wave_lds, s, bp);
      return hipGetLastError();
    }
  });
}

}  // namespace sqsg
