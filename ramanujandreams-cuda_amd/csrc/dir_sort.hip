// dir_sort.hip -- direction order for the fast walk kernels.
//
// A wave executes an axis-step body whenever one of its 64 lanes needs it,
// so lanes that walk different directions diverge: with random directions
// a wave can execute several times the axis steps any of its trajectories
// needs.  Before a large walk the trajectories are ordered by a 64-bit key
// of their direction (4 bits per axis) with a device radix sort; the
// kernels then map lane i to trajectory perm[i], so neighbouring lanes
// share their axis sequence.  Results are written at the trajectory's own
// index: the order is invisible to the caller.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include "walk_common.hpp"
#include "walk_kernels.hpp"

namespace cmfw {

__global__ void k_dir_keys(const int32_t *dir, int ld, int naxes, int n, uint64_t *keys,
                           int32_t *idx) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    uint64_t key = 0;
    for (int a = 0; a < naxes && a < 16; ++a)
        key |= (uint64_t)((dir[(size_t)a * ld + t] + 8) & 15) << (4 * (15 - a));
    keys[t] = key;
    idx[t] = t;
}

static size_t cub_temp_bytes(int n) {
    size_t b = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const uint64_t *)nullptr,
                                             (uint64_t *)nullptr, (const int32_t *)nullptr,
                                             (int32_t *)nullptr, n, 0, 64, (hipStream_t)0);
    return b;
}

size_t dir_sort_bytes(int n) {
    const size_t nn = (size_t)n;
    return 2 * align256(nn * 8) + 2 * align256(nn * 4) + align256(cub_temp_bytes(n));
}

hipError_t dir_sort(const int32_t *dir, int ld, int naxes, int n, void *ws, int32_t **perm,
                    hipStream_t st) {
    const size_t nn = (size_t)n;
    char *p = (char *)ws;
    uint64_t *k0 = (uint64_t *)p;
    p += align256(nn * 8);
    uint64_t *k1 = (uint64_t *)p;
    p += align256(nn * 8);
    int32_t *v0 = (int32_t *)p;
    p += align256(nn * 4);
    int32_t *v1 = (int32_t *)p;
    p += align256(nn * 4);
    size_t tb = cub_temp_bytes(n);
    hipLaunchKernelGGL(k_dir_keys, dim3((n + 255) / 256), dim3(256), 0, st, dir, ld, naxes, n, k0, v0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipcub::DeviceRadixSort::SortPairs((void *)p, tb, k0, k1, v0, v1, n, 0, 64, st);
    *perm = v1;
    return e;
}

}  // namespace cmfw
