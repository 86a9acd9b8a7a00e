// walk_screen.hip -- candidate screen for large searches (gfx950).
//
// Whether a walk fails is decided at three places that depend on the
// trajectory's position only, not on the walk matrix: the lead coefficient
// of the theta-companion column (theta_companion_column,
// cuda/cmf_pfq.cuh:127-181), the step denominator and, on inverse steps,
// the row-independent solve denominator (right_mul_affine_companion,
// cuda/cmf_pfq.cuh:186-239).  The screen evaluates exactly those, with the
// reference's operations in its order, over the first K trajectory steps
// and skips the O(d^2) matrix work.  A candidate it drops fails the real
// walk at that check or earlier; a candidate it keeps may still fail on a
// matrix-state check (non-finite or vanishing matrix), which the deep walk
// decides.  Start points off the integer grid (or beyond 2^23) are always
// kept.  One lane per candidate, no matrix in registers: high occupancy
// for the millions of candidates that die within a few axis steps.
#include "walk_kernels.hpp"
#include "walk_screen.hpp"

namespace cmfw {

template <int D>
__global__ __launch_bounds__(256) void k_screen(WalkArgs A, int K, int32_t *keep, int32_t *steps) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.n) return;
    int nf, ni;
    const bool ok = screen_traj<D>(A, t, K, nf, ni);
    keep[t] = ok ? 1 : 0;
    if (steps) steps[t] = nf + ni;
}

template <int D>
static hipError_t launch_screen_d(const WalkArgs &A, int K, int32_t *keep, int32_t *steps,
                                  hipStream_t st) {
    const int bs = 256;
    hipLaunchKernelGGL(k_screen<D>, dim3((A.n + bs - 1) / bs), dim3(bs), 0, st, A, K, keep, steps);
    return hipGetLastError();
}

hipError_t launch_screen(int D, const WalkArgs &A, int K, int32_t *keep, int32_t *steps,
                         hipStream_t st) {
    switch (D) {
        case 2: return launch_screen_d<2>(A, K, keep, steps, st);
        case 3: return launch_screen_d<3>(A, K, keep, steps, st);
        case 4: return launch_screen_d<4>(A, K, keep, steps, st);
        case 5: return launch_screen_d<5>(A, K, keep, steps, st);
        case 6: return launch_screen_d<6>(A, K, keep, steps, st);
        case 7: return launch_screen_d<7>(A, K, keep, steps, st);
        case 8: return launch_screen_d<8>(A, K, keep, steps, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cmfw
