// walk_screen.hpp -- the position-determined validity checks of a walk's
// first K trajectory steps (lead coefficient, step denominator, solve
// denominator), shared by the candidate screen and the mixed-batch engine.
#pragma once
#include "walk_common.hpp"

namespace cmfw {

// true when the first K steps of trajectory t pass every position check,
// evaluated with the reference's operations in its order (the checks of
// theta_companion_column and right_mul_affine_companion,
// cuda/cmf_pfq.cuh:146-235); nf / ni count the evaluated forward / inverse
// axis steps, the failing one included.  Start points off the integer grid
// (or beyond 2^23) pass with no step evaluated.
template <int D>
__device__ bool screen_traj(const WalkArgs &A, int t, int K, int &nf, int &ni) {
#pragma clang fp contract(off)
    int32_t dir[16];
    double xs[8], ys[8];
    load_traj(A, t, dir, xs, ys);
    nf = ni = 0;
    bool clean = true;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        clean = clean && xs[j] == rint(xs[j]) && fabs(xs[j]) < 8388608.0;
        clean = clean && ys[j] == rint(ys[j]) && fabs(ys[j]) < 8388608.0;
    }
    if (!clean) return true;
    int top = 0;
#pragma unroll
    for (int a = 0; a < 16; ++a) top = max(top, abs(dir[a]));
    bool ok = true;
    for (int step = 1; ok && step <= K; ++step) {
        for (int lev = top; ok && lev >= 1; --lev) {
            for (int ax = A.naxes - 1; ax >= 0; --ax) {
                const int dv = pick(dir, ax);
                if (abs(dv) < lev) continue;
                const int sign = dv > 0 ? 1 : -1;
                double xv[8], yv[8], denom;
                const bool inverse = axis_eval(xs, ys, A.p, ax, sign, xv, yv, denom);
                nf += inverse ? 0 : 1;
                ni += inverse ? 1 : 0;
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
    return ok;
}

}  // namespace cmfw
