// Alternative launch shapes of the dense search kernels (dense_kernels.hpp), which the
// autotuner times against a tile bound's default shape (dense.hip): the same tile bound with
// other tile groups or ring depths.  Shape 0 of a tile bound is its default; the shapes
// listed here are 1, 2, ...  They exist for jobs of one problem (split-mode jobs keep the
// default shape) and are instantiated in this compilation unit of their own.
#include "dense_kernels.hpp"

namespace sqsg {
namespace {

// f(shape) for alternative `alt` (>= 1) of a tile bound, hipErrorInvalidValue if it has none
template <class F> hipError_t with_alt_shape(uint32_t tile_bound, uint32_t alt, F &&f) {
  switch (tile_bound) {
    case 1:
      if (alt == 1) return f(DenseShape<1, 2, 2, 1>{});
      break;
    case 2:
      if (alt == 1) return f(DenseShape<2, 2, 2, 3>{});
      break;
    case 4:
      if (alt == 1) return f(DenseShape<4, 2, 2, 2>{});
      break;
    case 8:
      if (alt == 1) return f(DenseShape<8, 2, 2, 4>{});
      if (alt == 2) return f(DenseShape<8, 2, 2, 2>{});
      break;
    case 27:
      if (alt == 1) return f(DenseShape<27, 1, 2, 7>{});
      if (alt == 2) return f(DenseShape<27, 1, 2, 6>{});
      break;
    default: break;
  }
  return hipErrorInvalidValue;
}

}  // namespace

uint32_t dense_alt_shapes(uint32_t tile_bound) {
  uint32_t n = 0;
  while (with_alt_shape(tile_bound, n + 1, [](auto) { return hipSuccess; }) == hipSuccess) ++n;
  return n;
}

uint32_t dense_alt_groups(uint32_t tile_bound, uint32_t alt) {
  uint32_t g = 0;
  (void)with_alt_shape(tile_bound, alt, [&](auto sh) {
    g = decltype(sh)::G;
    return hipSuccess;
  });
  return g;
}

std::string dense_alt_shape_string(uint32_t tile_bound, uint32_t alt) {
  std::string s;
  (void)with_alt_shape(tile_bound, alt, [&](auto sh) {
    using Sh = decltype(sh);
    s = std::to_string(Sh::TT) + "," + std::to_string(Sh::G) + "," + std::to_string(Sh::W) +
        "," + std::to_string(Sh::D);
    return hipSuccess;
  });
  return s;
}

uint32_t dense_alt_waves_per_cu(uint32_t tile_bound, uint32_t alt, size_t wave_lds,
                                bool persistent) {
  int n = 0;
  const hipError_t e = with_alt_shape(tile_bound, alt, [&](auto sh) {
    using Sh = decltype(sh);
    const void *fn = persistent ? batch_dense_fn<Sh, false>() : run_dense_fn<Sh, false>();
    const hipError_t r = dense_lds(fn, wave_lds);
    if (r != hipSuccess) return r;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64, wave_lds);
  });
  return e == hipSuccess && n > 0 ? static_cast<uint32_t>(n) : 1u;
}

hipError_t launch_run_dense_alt(const RunParams &p, uint32_t tile_bound, uint32_t alt,
                                uint32_t blocks, hipStream_t s) {
  return with_alt_shape(tile_bound, alt, [&](auto sh) {
    using Sh = decltype(sh);
    const hipError_t e = dense_lds(run_dense_fn<Sh, false>(), p.wave_lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((run_dense_kernel<Sh::TT, Sh::G, Sh::W, Sh::D, false>), dim3(blocks),
                       dim3(64), p.wave_lds, s, p);
    return hipGetLastError();
  });
}

hipError_t prepare_batch_dense_alt(uint32_t tile_bound, uint32_t alt, size_t lds) {
  return with_alt_shape(tile_bound, alt, [&](auto sh) {
    return dense_lds(batch_dense_fn<decltype(sh), false>(), lds);
  });
}

hipError_t launch_batch_dense_alt(const BatchParams &bp, uint32_t tile_bound, uint32_t alt,
                                  uint32_t blocks, hipStream_t s) {
  return with_alt_shape(tile_bound, alt, [&](auto sh) {
    using Sh = decltype(sh);
    hipLaunchKernelGGL((batch_dense_kernel<Sh::TT, Sh::G, Sh::W, Sh::D, false>), dim3(blocks),
                       dim3(64), bp.wave_lds, s, bp);
    return hipGetLastError();
  });
}

}  // namespace sqsg
