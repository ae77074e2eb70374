// common.hpp - types and device helpers shared by the engine's kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "turboransac.h"

namespace tr {

constexpr int kTopK = 8;             // FP32 leaders re-scored in FP64
constexpr int kRrSlots = kTopK + 1;  // FP64 re-rank slots: the leaders and the initial pose
constexpr int kMaxSolutions = 4;  // P3P poses per minimal sample
constexpr int kPtBlock = 256;     // points per prepare/mask block
constexpr int kSampleBlock = 256; // samples per sampler block
constexpr int kPad = 256;         // FP32 point arrays are padded to this
constexpr int kMinChunk = 512;     // smallest points per scoring chunk a batch plan uses
constexpr int kLmThreads = 256;   // workgroup of the LM kernels, the same for every path
constexpr int kLmChunk = 2048;    // inliers per LM work item, the same for every path
constexpr int kLmSums = 28;       // cost, 21 JtJ (lower, row by row), 6 Jtr
constexpr int kLmMaxLaunches = 32;  // LM iteration launches at most
constexpr int kMaxDeviceRounds = 32;  // sampling rounds a captured batch launch holds
constexpr int kRoundsPerCheck = 8;    // rounds between host checks when a batch needs more
constexpr int kLmSoloMaxPoints = 4096;  // one frame per call up to this N: LM in one launch
// Launches of at most this many frames (one-frame calls, small batches,
// most served launches) refine frames of up to kLmSoloMaxPoints points in
// one launch, and the server scores them with the one-frame table's shape.
constexpr int kSmallLaunchFrames = 4;
constexpr int kLmServerBlocksPerCu = 8;  // LM iteration grid of the server's graphs, per CU
constexpr int kLmMaxFold = 8;       // chunk sums loaded at once by the folding workgroup

// ---------------------------------------------------------------------------
// Counter-based sampler.  Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// keyed by the frame seed; block b of the word stream of global sample g is
// Philox(ctr = (g_lo, g_hi, b, 0)).  Indices come from Lemire's
// multiply-shift with rejection and Floyd's 3-subset algorithm, so the
// stream is integer-only and identical on host and device.
// ---------------------------------------------------------------------------
struct Philox4 {
    uint32_t v[4];
};

__host__ __device__ inline Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                 uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Philox4{{c0, c1, c2, c3}};
}

constexpr int kMaxWordsPerSample = 64;

struct SampleStream {
    uint32_t k0, k1, g0, g1;
    uint32_t block;
    int used, drawn;
    Philox4 buf;
    __host__ __device__ SampleStream(uint64_t seed, uint64_t g)
        : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), g0((uint32_t)g),
          g1((uint32_t)(g >> 32)), block(0), used(4), drawn(0) {}
    __host__ __device__ uint32_t next() {
        if (used == 4) {
            buf = philox4x32_10(g0, g1, block, 0u, k0, k1);
            ++block;
            used = 0;
        }
        ++drawn;
        // Select without dynamic register indexing.
        const uint32_t w = used == 0 ? buf.v[0] : used == 1 ? buf.v[1] : used == 2 ? buf.v[2]
                                                                                  : buf.v[3];
        ++used;
        return w;
    }
    __host__ __device__ uint32_t bounded(uint32_t bound) {
        uint32_t x = next();
        uint64_t m = (uint64_t)x * bound;
        uint32_t lo = (uint32_t)m;
        if (lo < bound) {
            const uint32_t thr = (0u - bound) % bound;
            while (lo < thr && drawn < kMaxWordsPerSample) {
                x = next();
                m = (uint64_t)x * bound;
                lo = (uint32_t)m;
            }
        }
        return (uint32_t)(m >> 32);
    }
};

__host__ __device__ inline void sample_triple(uint64_t seed, uint32_t n, uint64_t g,
                                              int32_t out[3]) {
    SampleStream s(seed, g);
    const uint32_t i0 = s.bounded(n - 2u);
    uint32_t i1 = s.bounded(n - 1u);
    if (i1 == i0) i1 = n - 2u;
    uint32_t i2 = s.bounded(n);
    if (i2 == i0 || i2 == i1) i2 = n - 1u;
    out[0] = (int32_t)i0;
    out[1] = (int32_t)i1;
    out[2] = (int32_t)i2;
}

// Floyd's algorithm for a uniformly random 2-subset of [0, n) (PROSAC's
// samples that hold the newest point).
__host__ __device__ inline void sample_pair(uint64_t seed, uint32_t n, uint64_t g,
                                            int32_t out[2]) {
    SampleStream s(seed, g);
    const uint32_t i0 = s.bounded(n - 1u);
    uint32_t i1 = s.bounded(n);
    if (i1 == i0) i1 = n - 1u;
    out[0] = (int32_t)i0;
    out[1] = (int32_t)i1;
}

// Total order on floats as unsigned integers (for min-keys).
__host__ __device__ inline uint32_t float_order_key(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

}  // namespace tr
