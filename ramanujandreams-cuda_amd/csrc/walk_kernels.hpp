// walk_kernels.hpp -- host-side launchers of the walk kernels (one per TU).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cmfw {
struct WalkArgs;
hipError_t launch_walk_f64_ref(int D, const WalkArgs &A, hipStream_t st);
hipError_t launch_walk_f64_fast(int D, const WalkArgs &A, const int *bounds, int nseg,
                                void *seg_ws, hipStream_t st);
size_t fast_segment_bytes(int D, int n, int nseg);
hipError_t launch_walk_rns(int D, const WalkArgs &A, int np, uint32_t *resid, uint64_t *crt,
                           hipStream_t st);
}  // namespace cmfw
