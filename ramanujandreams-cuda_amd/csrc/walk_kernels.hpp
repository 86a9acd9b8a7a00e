// walk_kernels.hpp -- host-side launchers of the walk kernels (one per TU).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "walk_common.hpp"

namespace cmfw {
constexpr int kMaxRnsModuli = 64;
constexpr int kMaxRnsSegments = 256;

// The reference's modulus table (cuda/cmf_walk_rns.cu:26-43), host side.
inline uint32_t host_modulus(int k) {
    static const uint32_t t[kMaxRnsModuli] = {
        1073741789u, 1073741783u, 1073741741u, 1073741723u, 1073741717u, 1073741707u,
        1073741693u, 1073741683u, 1073741669u, 1073741657u, 1073741641u, 1073741623u,
        1073741611u, 1073741597u, 1073741587u, 1073741573u, 1073741563u, 1073741551u,
        1073741537u, 1073741527u, 1073741513u, 1073741503u, 1073741491u, 1073741477u,
        1073741467u, 1073741453u, 1073741441u, 1073741427u, 1073741413u, 1073741403u,
        1073741391u, 1073741377u, 1073741363u, 1073741353u, 1073741341u, 1073741329u,
        1073741317u, 1073741307u, 1073741293u, 1073741281u, 1073741267u, 1073741257u,
        1073741243u, 1073741231u, 1073741219u, 1073741207u, 1073741193u, 1073741183u,
        1073741171u, 1073741157u, 1073741147u, 1073741137u, 1073741123u, 1073741111u,
        1073741097u, 1073741087u, 1073741077u, 1073741063u, 1073741051u, 1073741039u,
        1073741027u, 1073741017u, 1073741003u, 1073740991u};
    return (k >= 0 && k < kMaxRnsModuli) ? t[k] : 0u;
}
hipError_t launch_walk_f64_ref(int D, const WalkArgs &A, hipStream_t st);
// ws: null, or screen_queue_bytes(n) bytes for the two-phase screen (which
// writes no step counts: steps must then be null)
hipError_t launch_screen(int D, const WalkArgs &A, int K, int32_t *keep, int32_t *steps, void *ws,
                         hipStream_t st);
size_t screen_queue_bytes(int n);
hipError_t launch_walk_f64_fast(int D, const WalkArgs &A, const int *bounds, int nseg,
                                void *seg_ws, const int32_t *perm, int layout, hipStream_t st);
size_t fast_segment_bytes(int D, int n, int nseg);
// direction order (dir_sort.hip): perm[i] = the i-th trajectory by direction key
size_t dir_sort_bytes(int n);
hipError_t dir_sort(const int32_t *dir, int ld, int naxes, int n, void *ws, int32_t **perm,
                    hipStream_t st);
hipError_t launch_walk_f32x2(int D, const WalkArgs &A, const int *bounds, int nseg, void *seg_ws,
                             const int32_t *perm, int layout, hipStream_t st);
hipError_t launch_walk_rns(int D, const WalkArgs &A, int np, uint32_t *resid, uint64_t *crt,
                           const int32_t *only, hipStream_t st);
size_t rns_seg_bytes(int D, int n, int np, int nseg);
int rns_plan(int N, int h, int n, int np, int64_t target_lanes, int min_steps, int *b,
             uint8_t *crt, int maxseg);
hipError_t launch_walk_rns_seg(int D, const WalkArgs &A, int np, const int *bounds,
                               const uint8_t *crtp, int nseg, void *ws, uint32_t *resid,
                               uint64_t *crt, int32_t **fallback, bool traj_major, bool reduce,
                               hipStream_t st);
// mixed batches (walk_batch.hip): one rank bucket's launch
struct BatchLaunch {
    WalkArgs w;              // shared inputs, snapshots, results
    const int32_t *bp, *bq;  // per trajectory p, q
    const double *bz;        // per trajectory z
    const int32_t *bN;       // per trajectory depth
    int h;
    const int32_t *queue;    // this bucket's trajectories, longest walk first
    int count;
    int32_t *counter;        // segment-item counter (zeroed before the segment kernel)
    void *ws;                // batch_bucket_bytes(D, count, maxseg)
    int target, min_steps;   // fast plan parameters
    int maxseg;              // most segments of a plan in this bucket
    int pp;                  // p of every trajectory (all pp F (pp-1)), or 0
};
size_t batch_bucket_bytes(int D, int count, int maxseg);
// exact verification walk (walk_exact_rns.hip)
int exact_primes(int K, uint32_t *out);
size_t exact_mods_bytes(int K);
void exact_fill_mods(const uint32_t *primes, int K, void *host_out);
hipError_t launch_walk_exact(int D, const WalkArgs &A, int K, int64_t zint, int zshift,
                             const void *mods, uint32_t *snap, int32_t *status, int32_t *fail_step,
                             hipStream_t st);
// part: 1 screen + compaction, 2 segment walk, 4 combine + reference-order walks
hipError_t launch_walk_batch(int D, const BatchLaunch &L, int block, int grid, int part,
                             hipStream_t st);
hipError_t launch_zero_i32(int32_t *p, int n, hipStream_t st);
// proposal network (mlp.hip)
size_t mlp_param_count();
size_t mlp_param_bf16_count();
size_t mlp_workspace_bytes(int B);
hipError_t mlp_forward(const float *P, const uint16_t *Pb, const uint16_t *X, int B, float *y,
                       void *ws, hipStream_t st);
hipError_t mlp_grad(const float *P, const uint16_t *Pb, const uint16_t *X, const float *target,
                    int B, float *grad, float *sq_out, void *ws, hipStream_t st);
hipError_t mlp_adam(float *P, uint16_t *Pb, const float *grad, float *m, float *v, int step,
                    float lr, hipStream_t st);
hipError_t mlp_sync_bf16(const float *P, uint16_t *Pb, hipStream_t st);
hipError_t mlp_features(const WalkArgs &A, int K, uint16_t *X, hipStream_t st);
hipError_t mlp_gemm_test(int mode, int M, int N, int K, const uint16_t *A, int lda,
                         const uint16_t *B, int ldb, float *C, int ldc, hipStream_t st);
}  // namespace cmfw
