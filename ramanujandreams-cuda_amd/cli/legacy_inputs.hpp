// legacy_inputs.hpp -- the inputs the reference drivers actually walk.
//
// The reference drivers generate start points and directions with
// std::mt19937_64 (cuda/cmf_walk_f64.cu:61-91) but write them into an array
// of 192-byte {int dir[16]; double v[16]} structs with the stride of the
// smaller record (128-byte points, cuda/cmf_walk_f64.cu:178; 16-int
// directions, :198), so values land in neighbouring structs.  A drop-in has
// to walk the same effective inputs to print the same JSON, so this header
// rebuilds that byte image and converts it to the engine's SoA layout.
// make_clean_inputs() gives what the generators meant (one record each).
#pragma once
#include <stdint.h>
#include <string.h>

#include <random>
#include <vector>

namespace cmfw_cli {

struct HostInputs {
    int n = 0, naxes = 0;
    std::vector<int32_t> dir;   // [naxes][n]
    std::vector<double> start;  // [naxes][n]
};

struct LegacyRecord {  // 192 bytes, the reference's TrajectoryInput
    int32_t dir[16];
    double v[16];
};
static_assert(sizeof(LegacyRecord) == 192, "record layout");

inline void to_soa(const std::vector<LegacyRecord> &img, int naxes, HostInputs &out) {
    const int n = (int)img.size();
    out.n = n;
    out.naxes = naxes;
    out.dir.assign((size_t)naxes * n, 0);
    out.start.assign((size_t)naxes * n, 0.0);
    for (int t = 0; t < n; ++t)
        for (int a = 0; a < naxes; ++a) {
            out.dir[(size_t)a * n + t] = img[t].dir[a];
            out.start[(size_t)a * n + t] = img[t].v[a];
        }
}

// mode 0 = limit, 1 = delta (cuda/cmf_walk_f64.cu:166-199)
inline void make_legacy_inputs(int p, int q, int n, int mode, bool zeta, uint64_t seed,
                               HostInputs &out) {
    const int naxes = p + q;
    std::vector<LegacyRecord> img((size_t)n);
    memset(img.data(), 0, img.size() * sizeof(LegacyRecord));
    if (mode == 0) {
        if (zeta) {
            for (int t = 0; t < n; ++t) {
                for (int a = 0; a < p; ++a) img[t].v[a] = 1.0;
                for (int a = 0; a < q; ++a) img[t].v[p + a] = 2.0;
            }
        } else {
            // 128-byte points written back to back over the 192-byte records
            unsigned char *bytes = reinterpret_cast<unsigned char *>(img.data());
            std::mt19937_64 rng(seed);
            for (int i = 0; i < n; ++i)
                for (int a = 0; a < naxes; ++a) {
                    double v = (double)(2 + (int)(rng() % 19));
                    memcpy(bytes + (size_t)i * 128 + (size_t)a * 8, &v, 8);
                }
        }
        for (int t = 0; t < n; ++t) {
            memset(img[t].dir, 0, sizeof(img[t].dir));
            img[t].dir[0] = 1;
        }
    } else {
        for (int t = 0; t < n; ++t)
            for (int a = 0; a < naxes; ++a)
                img[t].v[a] = zeta ? (a < p ? 1.0 : 2.0) : (a < p ? 2.0 : 3.0);
        // 16-int direction records written back to back over the 192-byte records
        unsigned char *bytes = reinterpret_cast<unsigned char *>(img.data());
        std::mt19937_64 rng(seed);
        const int dmax = 3;
        for (int i = 0; i < n; ++i) {
            int32_t d[16] = {0};
            bool zero = true;
            for (int a = 0; a < naxes; ++a) {
                d[a] = (int32_t)(rng() % (2 * dmax + 1)) - dmax;
                zero = zero && d[a] == 0;
            }
            if (zero) d[0] = 1;
            memcpy(bytes + (size_t)i * 64, d, (size_t)naxes * 4);
        }
    }
    to_soa(img, naxes, out);
}

// What the generators were meant to produce: point i in record i (limit
// mode), direction i in record i (delta mode).
inline void make_clean_inputs(int p, int q, int n, int mode, bool zeta, uint64_t seed,
                              HostInputs &out) {
    const int naxes = p + q;
    std::vector<LegacyRecord> img((size_t)n);
    memset(img.data(), 0, img.size() * sizeof(LegacyRecord));
    std::mt19937_64 rng(seed);
    for (int t = 0; t < n; ++t) {
        if (mode == 0) {
            for (int a = 0; a < naxes; ++a)
                img[t].v[a] = zeta ? (a < p ? 1.0 : 2.0) : (double)(2 + (int)(rng() % 19));
            img[t].dir[0] = 1;
        } else {
            for (int a = 0; a < naxes; ++a) img[t].v[a] = zeta ? (a < p ? 1.0 : 2.0) : (a < p ? 2.0 : 3.0);
            bool zero = true;
            for (int a = 0; a < naxes; ++a) {
                img[t].dir[a] = (int32_t)(rng() % 7) - 3;
                zero = zero && img[t].dir[a] == 0;
            }
            if (zero) img[t].dir[0] = 1;
        }
    }
    to_soa(img, naxes, out);
}

}  // namespace cmfw_cli
