// walk_f64.hip -- plain f64 CMF walk, reference operation order (gfx950).
//
// One trajectory per lane.  The D x D walk matrix (D = rank, a template
// parameter so every index is compile-time) stays in VGPRs for the whole
// walk; snapshots at N-2h, N-h, N go to HBM in SoA layout and are reduced to
// limit / delta / precision by the same lane afterwards.  The arithmetic is
// walk_exact.hpp's: bit-identical matrices to the CPU restatement.
#include "walk_exact.hpp"
#include "walk_kernels.hpp"

namespace cmfw {

// p known at compile time for the pFq family q = p - 1 (rank p or p - 1)
template <int D>
static hipError_t launch_ref(const WalkArgs &A, hipStream_t st) {
    const int bs = 256;
    const int grid = (A.n + bs - 1) / bs;
    constexpr int P1 = D + 1 <= 8 ? D + 1 : 0;
    if (A.q == A.p - 1 && A.p == D)
        hipLaunchKernelGGL((k_walk_f64_ref<D, D>), dim3(grid), dim3(bs), 0, st, A);
    else if (P1 > 0 && A.q == A.p - 1 && A.p == P1)
        hipLaunchKernelGGL((k_walk_f64_ref<D, P1>), dim3(grid), dim3(bs), 0, st, A);
    else
        hipLaunchKernelGGL((k_walk_f64_ref<D, 0>), dim3(grid), dim3(bs), 0, st, A);
    return hipGetLastError();
}

hipError_t launch_walk_f64_ref(int D, const WalkArgs &A, hipStream_t st) {
    switch (D) {
        case 2: return launch_ref<2>(A, st);
        case 3: return launch_ref<3>(A, st);
        case 4: return launch_ref<4>(A, st);
        case 5: return launch_ref<5>(A, st);
        case 6: return launch_ref<6>(A, st);
        case 7: return launch_ref<7>(A, st);
        case 8: return launch_ref<8>(A, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cmfw
