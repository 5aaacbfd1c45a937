// Runtime dispatch from the FFT size to the per-R kernel instantiations.
#include <hip/hip_runtime_api.h>

#include "xc_kernels.h"

namespace sxc {

hipError_t launch_forward(int R, const ForwardArgs& a, hipStream_t st) {
  if (a.n_rows <= 0) return hipSuccess;
  switch (R) {
    case 8: return launch_forward_r<8>(a, st);
    case 16: return launch_forward_r<16>(a, st);
    case 24: return launch_forward_r<24>(a, st);
    case 32: return launch_forward_r<32>(a, st);
    case 40: return launch_forward_r<40>(a, st);
    case 48: return launch_forward_r<48>(a, st);
    case 64: return launch_forward_r<64>(a, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_pairs(int R, bool full, const PairArgs& a, int grid, hipStream_t st) {
  if (a.n_work <= 0 || grid <= 0) return hipSuccess;
  switch (R) {
    case 8: return launch_pairs_r<8>(full, a, grid, st);
    case 16: return launch_pairs_r<16>(full, a, grid, st);
    case 24: return launch_pairs_r<24>(full, a, grid, st);
    case 32: return launch_pairs_r<32>(full, a, grid, st);
    case 40: return launch_pairs_r<40>(full, a, grid, st);
    case 48: return launch_pairs_r<48>(full, a, grid, st);
    case 64: return launch_pairs_r<64>(full, a, grid, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_resolve(int R, const ResolveArgs& a, int grid, hipStream_t st) {
  switch (R) {
    case 8: return launch_resolve_r<8>(a, grid, st);
    case 16: return launch_resolve_r<16>(a, grid, st);
    case 24: return launch_resolve_r<24>(a, grid, st);
    case 32: return launch_resolve_r<32>(a, grid, st);
    case 40: return launch_resolve_r<40>(a, grid, st);
    case 48: return launch_resolve_r<48>(a, grid, st);
    case 64: return launch_resolve_r<64>(a, grid, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace sxc
