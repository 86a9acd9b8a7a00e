// ref_shim.cpp -- host build of the reference's own cmf_pfq.cuh (TEST
// INFRASTRUCTURE ONLY; output goes to oracle/_ref/, never shipped).
//
// The header is included from where it lies under the reference tree (the
// include path is given by oracle/Makefile); nothing of it is copied here.
// The two CUDA function qualifiers it uses are defined empty, which is all
// a host compile of that header needs.  The per-trajectory tail below
// (result selection) restates walk_kernel_f64's body, cmf_walk_f64.cu:103-153,
// because the .cu file itself needs the CUDA runtime and cannot build here.
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#define __host__
#define __device__
#include "cmf_pfq.cuh"

extern "C" {

// One f64 trajectory through forward_walk + compute_limits + precision +
// kamidelta.  out6 = {limit, delta, p, q, precision, log_scale};
// snaps_out (nullable) = 3 x 64 doubles + 3 log scales.
int ref_walk_f64(int p, int q, double z, const int32_t *dir16, const double *start16, int N,
                 int h, double *out6, int32_t *n_limits, double *snaps_out) {
    CMFDesc desc = make_cmf_desc(p, q, z);
    int dir[MAX_AXES];
    Position start;
    for (int a = 0; a < MAX_AXES; ++a) {
        dir[a] = dir16[a];
        start.v[a] = start16[a];
    }
    WalkSnapshots snaps;
    memset(&snaps, 0, sizeof(snaps));
    bool ok = forward_walk(dir, start, desc, N, h, snaps);
    if (snaps_out) {
        const WalkMatrix *ms[3] = {&snaps.m1, &snaps.m2, &snaps.m3};
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 8; ++i)
                for (int j = 0; j < 8; ++j) snaps_out[k * 64 + i * 8 + j] = ms[k]->a[i][j];
        snaps_out[192] = snaps.log_scale1;
        snaps_out[193] = snaps.log_scale2;
        snaps_out[194] = snaps.log_scale3;
    }
    if (!ok) {
        for (int k = 0; k < 6; ++k) out6[k] = NAN;
        *n_limits = 0;
        return 0;
    }
    static thread_local LimitResult lim[256];
    int nl = 0;
    compute_limits(snaps, lim, nl, 256);
    double L = NAN;
    for (int i = 0; i < nl; ++i)
        if (lim[i].method == 1) { L = lim[i].L; break; }
    if (isnan(L) && nl > 0) L = lim[0].L;
    const int col = desc.rank - 1;
    out6[0] = L;
    out6[2] = snaps.m3.a[0][col];
    out6[3] = snaps.m3.a[1][col];
    out6[4] = compute_precision(snaps);
    out6[5] = snaps.log_scale3;
    out6[1] = compute_kamidelta(snaps, L);
    *n_limits = nl;
    return 1;
}

// Batch over trajectory-major inputs [n][16]; OpenMP over trajectories.
void ref_walk_f64_batch(int p, int q, double z, int n, const int32_t *dirs,
                        const double *starts, int N, int h, double *out6,
                        int32_t *n_limits, int32_t *valid, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int i = 0; i < n; ++i)
        valid[i] = ref_walk_f64(p, q, z, dirs + (size_t)i * 16, starts + (size_t)i * 16, N, h,
                                out6 + (size_t)i * 6, n_limits + i, nullptr);
}

}  // extern "C"
