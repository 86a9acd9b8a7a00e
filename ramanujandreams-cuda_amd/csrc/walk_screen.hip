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
#include "walk_common.hpp"
#include "walk_kernels.hpp"

namespace cmfw {

template <int D>
__global__ __launch_bounds__(256) void k_screen(WalkArgs A, int K, int32_t *keep, int32_t *steps) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.n) return;
    int32_t dir[16];
    double xs[8], ys[8];
    load_traj(A, t, dir, xs, ys);
    bool clean = true;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        clean = clean && xs[j] == rint(xs[j]) && fabs(xs[j]) < 8388608.0;
        clean = clean && ys[j] == rint(ys[j]) && fabs(ys[j]) < 8388608.0;
    }
    if (!clean) {
        keep[t] = 1;
        if (steps) steps[t] = 0;
        return;
    }
    int top = 0;
#pragma unroll
    for (int a = 0; a < 16; ++a) top = max(top, abs(dir[a]));
    bool ok = true;
    int evaluated = 0;
    for (int step = 1; ok && step <= K; ++step) {
        for (int lev = top; ok && lev >= 1; --lev) {
            for (int ax = A.naxes - 1; ax >= 0; --ax) {
                const int dv = pick(dir, ax);
                if (abs(dv) < lev) continue;
                const int sign = dv > 0 ? 1 : -1;
                double xv[8], yv[8], denom;
                const bool inverse = axis_eval(xs, ys, A.p, ax, sign, xv, yv, denom);
                ++evaluated;
                double col[D];
                ok = companion_column_ref<D>(xv, yv, A.p, A.q, A.z, col) && isfinite(denom) &&
                     !(fabs(denom) <= 64.0 * kEps);
                if (ok && inverse) {
                    const double ia = 1.0 / denom;
                    double be[D];
                    be[D - 1] = 1.0;
#pragma unroll
                    for (int j = D - 2; j >= 0; --j) be[j] = -be[j + 1] * ia;
                    double db = 0.0;
#pragma unroll
                    for (int k = 0; k < D; ++k) db = db + be[k] * col[k];
                    const double sd = 1.0 + db * ia;
                    const double ss = fmax(1.0, fabs(db * ia));
                    ok = isfinite(sd) && !(fabs(sd) <= 128.0 * kEps * ss);
                }
                if (!ok) break;
                axis_move(xs, ys, A.p, ax, sign);
            }
        }
    }
    keep[t] = ok ? 1 : 0;
    if (steps) steps[t] = evaluated;
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
