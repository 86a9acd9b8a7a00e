// walk_batch.hip -- device side of mixed batches: trajectories of different
// CMF (p, q), z and depth N walked in one submission.
//
// The host planner (batch.hip) puts the trajectories of each rank D in a
// device work queue, longest walks first.  k_walk_batch<D> is a persistent
// kernel: each wave takes the next tile of 64 queue entries with one atomic
// on the rank's counter and walks them, one trajectory per lane, in the
// fast FP64 engine (walk_full_fast: no depth segments, the batch fills the
// GPU by itself).  Longest-first order keeps the lanes of a wave on walks of
// similar length and leaves the short walks to fill the tail.  Trajectories
// the fast engine cannot decide are flagged and walked in the reference's
// order by k_walk_batch_exact<D>, which carries that path's register
// footprint alone.
#include "walk_fast.hpp"
#include "walk_kernels.hpp"

namespace cmfw {

struct BatchArgs {
    WalkArgs w;              // dir, start, ld, snapshots (laid out for rank D), results
    const int32_t *bp, *bq;  // per trajectory p, q
    const double *bz;        // per trajectory z
    const int32_t *bN;       // per trajectory depth
    int h;
    const int32_t *queue;    // trajectory indices of this rank
    int count;
    int32_t *counter;        // tile counter, zero at launch
    int32_t *exact;          // [n] 1 where the reference order must walk
};

__device__ __forceinline__ WalkArgs traj_args(const BatchArgs &B, int t) {
    WalkArgs A = B.w;
    A.p = B.bp[t];
    A.q = B.bq[t];
    A.naxes = A.p + A.q;
    A.z = B.bz[t];
    A.N = B.bN[t];
    A.h = B.h;
    return A;
}

template <int D>
__global__ __launch_bounds__(256) void k_walk_batch(BatchArgs B) {
    const int lane = threadIdx.x & 63;
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(B.counter, 64);
        base = __shfl(base, 0, 64);
        if (base >= B.count) break;  // every wave reaches this once the queue is drained
        const int i = base + lane;
        if (i < B.count) {
            const int t = B.queue[i];
            const WalkArgs A = traj_args(B, t);
            B.exact[t] = walk_full_fast<D, ArithF64>(A, t) ? 0 : 1;
        }
    }
}

template <int D>
__global__ __launch_bounds__(256) void k_walk_batch_exact(BatchArgs B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.count) return;
    const int t = B.queue[i];
    if (!B.exact[t]) return;
    walk_exact_traj<D>(traj_args(B, t), t);
}

__global__ void k_zero_i32(int32_t *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <int D>
static hipError_t launch_batch_d(const BatchArgs &B, int block, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_walk_batch<D>, dim3(grid), dim3(block), 0, st, B);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_walk_batch_exact<D>, dim3((B.count + 255) / 256), dim3(256), 0, st, B);
    return hipGetLastError();
}

hipError_t launch_walk_batch(int D, const WalkArgs &w, const int32_t *bp, const int32_t *bq,
                             const double *bz, const int32_t *bN, int h, const int32_t *queue,
                             int count, int32_t *counter, int32_t *exact, int block, int grid,
                             hipStream_t st) {
    BatchArgs B;
    B.w = w;
    B.bp = bp;
    B.bq = bq;
    B.bz = bz;
    B.bN = bN;
    B.h = h;
    B.queue = queue;
    B.count = count;
    B.counter = counter;
    B.exact = exact;
    switch (D) {
        case 2: return launch_batch_d<2>(B, block, grid, st);
        case 3: return launch_batch_d<3>(B, block, grid, st);
        case 4: return launch_batch_d<4>(B, block, grid, st);
        case 5: return launch_batch_d<5>(B, block, grid, st);
        case 6: return launch_batch_d<6>(B, block, grid, st);
        case 7: return launch_batch_d<7>(B, block, grid, st);
        case 8: return launch_batch_d<8>(B, block, grid, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_zero_i32(int32_t *p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace cmfw
