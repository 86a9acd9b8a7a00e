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
#include <algorithm>

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

// Two-phase screen.  Phase 1 (k_screen_pre): the exact integer checks of
// screen_int for every candidate, a few integer operations per axis step;
// the candidates they cannot decide are appended to a queue (one atomic per
// block).  Phase 2 (k_screen_queue): screen_traj for the queued candidates
// only.  Most candidates fail an integer check, so the FP64 column work
// runs on a small fraction of them, in waves that are all busy.  keep[] is
// indexed by candidate, so the result does not depend on the queue order
// and equals k_screen's.
// One atomic per block of 1024 (the waves' counts are summed in LDS):
// tens of thousands of same-address atomics, one per wave, serialise in the
// L2 and cost more than the integer checks themselves.
template <int NAX>
__global__ __launch_bounds__(1024) void k_screen_pre(WalkArgs A, int rank, int K, int32_t *keep,
                                                     int32_t *queue, int32_t *count) {
    __shared__ int wcount[16], wbase[16];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int r = 0;
    if (t < A.n) {
        r = screen_int<NAX>(A, rank, t, K);
        keep[t] = (r == 0) ? 0 : 1;  // undecided ones are rewritten by phase 2
    }
    const bool need = r == 1;
    const unsigned long long m = __ballot(need);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wcount[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        int tot = 0;
        for (int w = 0; w < nw; ++w) {
            wbase[w] = tot;
            tot += wcount[w];
        }
        const int base = tot > 0 ? atomicAdd(count, tot) : 0;
        for (int w = 0; w < nw; ++w) wbase[w] += base;
    }
    __syncthreads();
    if (need) queue[wbase[wave] + __popcll(m & ((1ull << lane) - 1ull))] = t;
}

template <int D, int PP>
__global__ __launch_bounds__(256) void k_screen_queue(WalkArgs A, int K, int32_t *keep,
                                                      const int32_t *queue, const int32_t *count,
                                                      bool fwd_decided) {
    const int c = *count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x) {
        const int t = queue[i];
        int nf, ni;
        keep[t] = screen_traj<D, PP>(A, t, K, nf, ni, fwd_decided) ? 1 : 0;
    }
}

size_t screen_queue_bytes(int n) { return align256((size_t)n * 4) + 256; }

template <int D>
static hipError_t launch_screen_d(const WalkArgs &A, int K, int32_t *keep, int32_t *steps,
                                  void *ws, hipStream_t st) {
    const int bs = 256;
    const int g = (A.n + bs - 1) / bs;
    if (!ws) {
        hipLaunchKernelGGL(k_screen<D>, dim3(g), dim3(bs), 0, st, A, K, keep, steps);
        return hipGetLastError();
    }
    int32_t *queue = (int32_t *)ws;
    int32_t *count = (int32_t *)((char *)ws + align256((size_t)A.n * 4));
    hipError_t e = launch_zero_i32(count, 1, st);
    if (e != hipSuccess) return e;
    hipEvent_t t0 = kernel_timer_begin(st);
    const int gp = (A.n + 1023) / 1024;
    switch (A.naxes) {
#define CMFW_PRE(NAX)                                                                            \
    case NAX:                                                                                    \
        hipLaunchKernelGGL(k_screen_pre<NAX>, dim3(gp), dim3(1024), 0, st, A, D, K, keep, queue, \
                           count);                                                               \
        break;
        CMFW_PRE(2) CMFW_PRE(3) CMFW_PRE(4) CMFW_PRE(5) CMFW_PRE(6) CMFW_PRE(7) CMFW_PRE(8)
        CMFW_PRE(9) CMFW_PRE(10) CMFW_PRE(11) CMFW_PRE(12) CMFW_PRE(13) CMFW_PRE(14) CMFW_PRE(15)
        CMFW_PRE(16)
#undef CMFW_PRE
        default: return hipErrorInvalidValue;
    }
    kernel_timer_end("screen_pre", t0, st);
    // phase 2 grid: enough lanes for a queue of a tenth of the candidates in
    // one pass, strided beyond that
    const int g2 = std::max(1, std::min(g, std::max(1024, g / 10)));
    const bool fwd = screen_fwd_decided(A.p, A.q, A.z, D);
    // p known at compile time for the q = p - 1 family
    constexpr int P1 = D + 1 <= 8 ? D + 1 : 0;
    if (A.q == A.p - 1 && A.p == D)
        hipLaunchKernelGGL((k_screen_queue<D, D>), dim3(g2), dim3(bs), 0, st, A, K, keep,
                           (const int32_t *)queue, (const int32_t *)count, fwd);
    else if (P1 > 0 && A.q == A.p - 1 && A.p == P1)
        hipLaunchKernelGGL((k_screen_queue<D, P1>), dim3(g2), dim3(bs), 0, st, A, K, keep,
                           (const int32_t *)queue, (const int32_t *)count, fwd);
    else
        hipLaunchKernelGGL((k_screen_queue<D, 0>), dim3(g2), dim3(bs), 0, st, A, K, keep,
                           (const int32_t *)queue, (const int32_t *)count, fwd);
    return hipGetLastError();
}

hipError_t launch_screen(int D, const WalkArgs &A, int K, int32_t *keep, int32_t *steps, void *ws,
                         hipStream_t st) {
    switch (D) {
        case 2: return launch_screen_d<2>(A, K, keep, steps, ws, st);
        case 3: return launch_screen_d<3>(A, K, keep, steps, ws, st);
        case 4: return launch_screen_d<4>(A, K, keep, steps, ws, st);
        case 5: return launch_screen_d<5>(A, K, keep, steps, ws, st);
        case 6: return launch_screen_d<6>(A, K, keep, steps, ws, st);
        case 7: return launch_screen_d<7>(A, K, keep, steps, ws, st);
        case 8: return launch_screen_d<8>(A, K, keep, steps, ws, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cmfw
