// walk_f64_fast.hip -- the f64 CMF walk engineered for the CDNA4 VALU.
//
// Same recurrence and validity rules as the reference-order walk, with:
//  * depth segments: W_N = S_1 S_2 ... S_K where S_k is the product of the
//    axis matrices of trajectory steps (b_{k-1}, b_k].  Its start point is
//    start + b_{k-1} * dir, so every (trajectory, segment) pair is an
//    independent lane; a second kernel multiplies the K segment matrices
//    in order and takes the snapshots at N-2h, N-h, N (segment boundaries
//    always include them).  A few thousand deep trajectories then fill the
//    256 CUs instead of a few dozen waves.
//  * FMA arithmetic; the theta polynomials are rebuilt per axis step with
//    fused multiply-adds; 1/lead and 1/solve-denominator are computed once.
//  * exact power-of-two rescaling instead of the per-step max/divide/log of
//    normalize_projective: the max is still taken every axis step (one
//    v_max per entry), but the matrix is rescaled only when the max leaves
//    [2^-400, 2^400]; the scale is an integer exponent and its log is taken
//    once per snapshot.
//  * the solve denominator of an inverse step is row-independent and
//    computed once per step (the reference recomputes it per row).
// Validity parity: every failure the reference can raise depends on the
// position only (lead, denominator, solve denominator).  lead and the
// denominator are exact integers here; a solve denominator that falls
// within a rounding-error bound of the threshold is re-evaluated in the
// reference's operation order before deciding.  Trajectories whose start
// is not an integer point below 2^30 in magnitude take the reference-order
// walk (walk_exact.hpp) instead.
#include "walk_exact.hpp"
#include "walk_kernels.hpp"

namespace cmfw {

struct SegArgs {
    WalkArgs w;
    int nseg;
    int b[kMaxSegments + 1];  // segment boundaries in trajectory steps, b[0]=0, b[nseg]=N
    double *segM;             // [nseg][D][D][n]
    int32_t *segE;            // [nseg][n] power-of-two exponent of each segment product
    int32_t *segV;            // [nseg][n] 1 ok, 0 failed, 2 exact path
    int32_t *segF, *segI;     // [nseg][n] executed fwd / inv axis steps
};

__device__ __forceinline__ bool clean_point(double v) {
    return v == rint(v) && fabs(v) < 1073741824.0;
}

// theta-companion column with fused multiply-adds (same polynomial as
// companion_column_ref); returns false on a vanishing lead.
template <int D>
__device__ __forceinline__ bool companion_column_fast(const double (&xv)[8], const double (&yv)[8],
                                                      int p, int q, double z, double (&col)[D],
                                                      double &csum) {
    double px[9], py[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        px[k] = 0.0;
        py[k] = 0.0;
    }
    px[0] = 1.0;
    py[0] = 1.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j < p) {
#pragma unroll
            for (int k = j + 1; k >= 1; --k) px[k] = fma(px[k], xv[j], px[k - 1]);
            px[0] = px[0] * xv[j];
        }
        if (j < q) {
#pragma unroll
            for (int k = j + 1; k >= 1; --k) py[k] = fma(py[k], yv[j], py[k - 1]);
            py[0] = py[0] * yv[j];
        }
    }
    double dc[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) dc[k] = 0.0;
#pragma unroll
    for (int k = 0; k <= 8; ++k)
        if (k <= p) dc[k] = -z * px[k];
#pragma unroll
    for (int k = 0; k <= 8; ++k)
        if (k <= q) dc[k + 1] = dc[k + 1] + py[k];
    // lead of D(theta) and the reference's vanishing test; for an integer
    // point both are exact, so the decision equals the reference's
    const double lead = dc[D];
    double lscale = 1.0;
    if (D == p && p == q + 1) {
        lscale = fmax(1.0, fabs(1.0 - z));
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (i < p) lscale += fabs(xv[i]);
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (i < q) lscale += fabs(yv[i]);
    }
    if (!isfinite(lead) || fabs(lead) <= 64.0 * kEps * lscale) return false;
    const double nil = -1.0 / lead;
    csum = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        col[i] = dc[i] * nil;
        csum += fabs(col[i]);
    }
    return isfinite(csum);
}

// exact rescale of M by a power of two so that max |M| lies in [0.5, 1)
template <int D>
__device__ __forceinline__ void rescale_pow2(double (&M)[D][D], int &e2) {
    double mx = 0.0;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) mx = fmax(mx, fabs(M[r][c]));
    if (!(mx > 0.0) || !isfinite(mx)) return;
    const int e = ilogb(mx) + 1;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) M[r][c] = ldexp(M[r][c], -e);
    e2 += e;
}

// Keeps max |M| inside [2^-400, 2^400] with exact power-of-two rescales,
// and fails as normalize_projective does (cuda/cmf_pfq.cuh:109) when the
// max is zero or not finite; fmax skips NaN entries as the reference's
// max does.  One v_max per entry per axis step; the rescale itself is rare.
template <int D>
__device__ __forceinline__ bool keep_in_range(double (&M)[D][D], int &e2) {
    double mx = 0.0;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) mx = fmax(mx, fabs(M[r][c]));
    if (!(mx > 0.0) || !isfinite(mx)) return false;
    if (mx > 0x1p400 || mx < 0x1p-400) {
        const int e = ilogb(mx) + 1;
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int c = 0; c < D; ++c) M[r][c] = ldexp(M[r][c], -e);
        e2 += e;
    }
    return true;
}

// One axis step on the fast path.
template <int D>
__device__ __forceinline__ bool axis_step_fast(double (&M)[D][D], const double (&xs)[8],
                                               const double (&ys)[8], int p, int q, double z,
                                               int ax, int sign, bool &inverse) {
    double xv[8], yv[8], denom;
    inverse = axis_eval(xs, ys, p, ax, sign, xv, yv, denom);
    double col[D], csum;
    if (!companion_column_fast<D>(xv, yv, p, q, z, col, csum)) return false;
    if (denom == 0.0) return false;  // |denom| <= 64 eps for an integer point
    const double ia = 1.0 / denom;
    const double aia = fabs(ia);
    if (!inverse) {
#pragma unroll
        for (int r = 0; r < D; ++r) {
            double dot = M[r][0] * col[0];
#pragma unroll
            for (int k = 1; k < D; ++k) dot = fma(M[r][k], col[k], dot);
#pragma unroll
            for (int j = 0; j + 1 < D; ++j) M[r][j] = fma(M[r][j + 1], ia, M[r][j]);
            M[r][D - 1] = fma(dot, ia, M[r][D - 1]);
        }
        return true;
    }
    double be[D];
    be[D - 1] = 1.0;
#pragma unroll
    for (int j = D - 2; j >= 0; --j) be[j] = -be[j + 1] * ia;
    double db = be[0] * col[0], dabs = fabs(be[0] * col[0]);
#pragma unroll
    for (int k = 1; k < D; ++k) {
        db = fma(be[k], col[k], db);
        dabs += fabs(be[k] * col[k]);
    }
    double sd = fma(db, ia, 1.0);
    // The fast sd decides only when it is far from the threshold
    // 128 eps max(1, |db/a|): more than 2^20 times the size of its own
    // terms' rounding.  A negative evaluation coordinate can make the
    // theta polynomials cancel, so it always takes the reference's path.
    bool neg = false;
#pragma unroll
    for (int j = 0; j < 8; ++j) neg = neg || (j < p && xv[j] < 0.0) || (j < q && yv[j] < 0.0);
    if (neg || !(fabs(sd) > 0x1p-20 * (dabs * aia + 1.0))) {
        // near the threshold: decide exactly as the reference does
        double cref[D];
        if (!companion_column_ref<D>(xv, yv, p, q, z, cref)) return false;
        // (this file is compiled without implicit contraction: the
        // products and sums below round separately, as in the reference)
        double bref[D];
        bref[D - 1] = 1.0;
        double dbr = 0.0;
#pragma unroll
        for (int j = D - 2; j >= 0; --j) bref[j] = -bref[j + 1] * ia;
#pragma unroll
        for (int k = 0; k < D; ++k) dbr = dbr + bref[k] * cref[k];
        const double sdr = 1.0 + dbr * ia;
        const double ssr = fmax(1.0, fabs(dbr * ia));
        if (!isfinite(sdr) || fabs(sdr) <= 128.0 * kEps * ssr) return false;
        sd = sdr;
#pragma unroll
        for (int i = 0; i < D; ++i) col[i] = cref[i];
    }
    if (!isfinite(sd)) return false;
    const double isd = 1.0 / sd;
#pragma unroll
    for (int r = 0; r < D; ++r) {
        double al[D];
        al[D - 1] = 0.0;
        double da = 0.0;
#pragma unroll
        for (int j = D - 2; j >= 0; --j) al[j] = fma(-al[j + 1], ia, M[r][j]);
#pragma unroll
        for (int k = 0; k < D - 1; ++k) da = fma(al[k], col[k], da);
        const double tt = fma(-da, ia, M[r][D - 1]) * isd;
#pragma unroll
        for (int j = 0; j < D; ++j) M[r][j] = fma(be[j], tt, al[j]);
    }
    return true;
}

template <int D>
__global__ __launch_bounds__(256) void k_walk_seg(SegArgs S) {
    const WalkArgs &A = S.w;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (t >= A.n) return;
    const size_t o = (size_t)s * A.n + t;
    int32_t dir[16];
    double xs[8], ys[8];
    load_traj(A, t, dir, xs, ys);
    bool clean = true;
#pragma unroll
    for (int j = 0; j < 8; ++j) clean = clean && clean_point(xs[j]) && clean_point(ys[j]);
    if (!clean) {
        S.segV[o] = 2;
        S.segF[o] = 0;
        S.segI[o] = 0;
        return;
    }
    // start of this segment: start + b_s * dir (exact integers)
    const double b0 = (double)S.b[s];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        xs[j] = fma(b0, (double)(j < A.p ? dir[j] : 0), xs[j]);
        ys[j] = fma(b0, (double)(j < A.q ? pick(dir, A.p + j) : 0), ys[j]);
    }
    double M[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) M[i][j] = (i == j) ? 1.0 : 0.0;
    int top = 0;
#pragma unroll
    for (int a = 0; a < 16; ++a) top = max(top, abs(dir[a]));
    int e2 = 0, nf = 0, ni = 0;
    bool ok = true;
    const int steps = S.b[s + 1] - S.b[s];
    for (int st = 0; ok && st < steps; ++st) {
        for (int lev = top; ok && lev >= 1; --lev) {
            for (int ax = A.naxes - 1; ax >= 0; --ax) {
                const int dv = pick(dir, ax);
                if (abs(dv) < lev) continue;
                const int sign = dv > 0 ? 1 : -1;
                bool inverse;
                ok = axis_step_fast<D>(M, xs, ys, A.p, A.q, A.z, ax, sign, inverse) &&
                     keep_in_range<D>(M, e2);
                nf += inverse ? 0 : 1;
                ni += inverse ? 1 : 0;
                if (!ok) break;
                axis_move(xs, ys, A.p, ax, sign);
            }
        }
    }
    if (ok) rescale_pow2<D>(M, e2);
    S.segF[o] = nf;
    S.segI[o] = ni;
    S.segV[o] = ok ? 1 : 0;
    if (!ok) return;
    S.segE[o] = e2;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) S.segM[((size_t)((s * D + i) * D + j)) * A.n + t] = M[i][j];
}

template <int D>
__global__ __launch_bounds__(256) void k_walk_combine(SegArgs S) {
    const WalkArgs &A = S.w;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.n) return;
    if (S.segV[t] == 2) {  // not an integer point: reference-order walk
        walk_exact_traj<D>(A, t);
        return;
    }
    const int N = A.N, h = A.h;
    const int cp1 = N - 2 * h, cp2 = N - h;
    double W[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) W[i][j] = (i == j) ? 1.0 : 0.0;
    int e2 = 0, nf = 0, ni = 0;
    bool ok = !(N <= 0 || h <= 0 || N - 2 * h <= 0);
    for (int s = 0; ok && s < S.nseg; ++s) {
        const size_t o = (size_t)s * A.n + t;
        nf += S.segF[o];
        ni += S.segI[o];
        if (S.segV[o] != 1) {
            ok = false;
            break;
        }
        double G[D][D];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) G[i][j] = S.segM[((size_t)((s * D + i) * D + j)) * A.n + t];
        // W <- W * G, one row at a time
#pragma unroll
        for (int r = 0; r < D; ++r) {
            double row[D];
#pragma unroll
            for (int j = 0; j < D; ++j) {
                double acc = W[r][0] * G[0][j];
#pragma unroll
                for (int k = 1; k < D; ++k) acc = fma(W[r][k], G[k][j], acc);
                row[j] = acc;
            }
#pragma unroll
            for (int j = 0; j < D; ++j) W[r][j] = row[j];
        }
        e2 += S.segE[o];
        rescale_pow2<D>(W, e2);
        const int bend = S.b[s + 1];
        for (int k = 0; k < 3; ++k) {
            const int cp = (k == 0) ? cp1 : (k == 1) ? cp2 : N;
            if (cp != bend) continue;
            // the reference's snapshot: max-normalised, log of the scale
            double mx = 0.0;
#pragma unroll
            for (int i = 0; i < D; ++i)
#pragma unroll
                for (int j = 0; j < D; ++j) mx = fmax(mx, fabs(W[i][j]));
            if (!(mx > 0.0) || !isfinite(mx)) {
                ok = false;
                break;
            }
            const double inv = 1.0 / mx;
#pragma unroll
            for (int i = 0; i < D; ++i)
#pragma unroll
                for (int j = 0; j < D; ++j)
                    A.snap[((size_t)((k * D + i) * D + j)) * A.sld + t] = W[i][j] * inv;
            A.snap_ls[(size_t)k * A.sld + t] = (double)e2 * 0.6931471805599453 + log(mx);
        }
    }
    if (A.n_fwd) A.n_fwd[t] = nf;
    if (A.n_inv) A.n_inv[t] = ni;
    finish_from_snapshots<D>(A, t, ok);
}

template <int D>
static hipError_t launch_fast_d(const SegArgs &S, hipStream_t st) {
    const int bs = 256;
    const int gx = (S.w.n + bs - 1) / bs;
    hipLaunchKernelGGL(k_walk_seg<D>, dim3(gx, S.nseg), dim3(bs), 0, st, S);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_walk_combine<D>, dim3(gx), dim3(bs), 0, st, S);
    return hipGetLastError();
}

hipError_t launch_walk_f64_fast(int D, const WalkArgs &A, const int *bounds, int nseg,
                                void *seg_ws, hipStream_t st) {
    SegArgs S;
    S.w = A;
    S.nseg = nseg;
    for (int i = 0; i <= nseg; ++i) S.b[i] = bounds[i];
    char *p = (char *)seg_ws;
    const size_t n = (size_t)A.n;
    S.segM = (double *)p;
    p += align256((size_t)nseg * D * D * n * sizeof(double));
    S.segE = (int32_t *)p;
    p += align256((size_t)nseg * n * 4);
    S.segV = (int32_t *)p;
    p += align256((size_t)nseg * n * 4);
    S.segF = (int32_t *)p;
    p += align256((size_t)nseg * n * 4);
    S.segI = (int32_t *)p;
    switch (D) {
        case 2: return launch_fast_d<2>(S, st);
        case 3: return launch_fast_d<3>(S, st);
        case 4: return launch_fast_d<4>(S, st);
        case 5: return launch_fast_d<5>(S, st);
        case 6: return launch_fast_d<6>(S, st);
        case 7: return launch_fast_d<7>(S, st);
        case 8: return launch_fast_d<8>(S, st);
        default: return hipErrorInvalidValue;
    }
}

size_t fast_segment_bytes(int D, int n, int nseg) {
    const size_t nn = (size_t)n;
    return align256((size_t)nseg * D * D * nn * sizeof(double)) + 4 * align256((size_t)nseg * nn * 4);
}

}  // namespace cmfw
