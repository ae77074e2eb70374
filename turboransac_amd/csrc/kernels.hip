// kernels.hip - the pipeline's HIP kernels for gfx950 (CDNA4, wave64).
//
// Stage map to the reference (Module/TbrCudaPnP):
//   k_prepare     camera.unproject + FP32 copies      entry.cu:42-49
//   k_sample_p3p  sampling + P3P per minimal sample   ransac.cuh:176-241, p3p.cuh
//   k_msac*       MSAC residuals over H x N           ransac.cuh:259-335, msac.cuh:18-56
//   k_select      per-frame best candidates           ransac.cuh:87-137, turbo_ransac.cu:185-221
//   k_rerank      FP64 MSAC of the FP32 leaders       msac.cuh:18-56 (FP64 form)
//   k_choose      best hypothesis, running over rounds turbo_ransac.cu:235-251
//   k_mask        inlier mask of the RANSAC pose      entry.cu:12-26
//   k_lm_compact  inlier compaction, LM start         entry.cu:87-132
//   k_lm_iter     Levenberg-Marquardt iterations      lm.cuh:15-171, jaco.cuh
//
// This file is compiled with -ffp-contract=off: the FP64 stages (P3P,
// re-rank, mask, LM) round operation by operation like the host oracle.  The
// FP32 scoring kernel spells its fused multiply-adds out with fmaf.
#include <float.h>

#include "kernels.hpp"
#include "p3p.hpp"

namespace tr {

// ---------------------------------------------------------------------------
// Wave-level helpers (wave64).
// ---------------------------------------------------------------------------
__device__ inline uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

__device__ inline double wave_sum_f64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Exclusive prefix sum over the wave of a value in [0, 7], from ballots.
__device__ inline uint32_t wave_prefix_small(uint32_t v, uint32_t* total) {
    const uint64_t lt = (threadIdx.x & 63) ? (~0ull >> (64 - (threadIdx.x & 63))) : 0ull;
    const uint64_t b0 = __ballot(v & 1u), b1 = __ballot(v & 2u), b2 = __ballot(v & 4u);
    *total = (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1) + 4u * (uint32_t)__popcll(b2);
    return (uint32_t)__popcll(b0 & lt) + 2u * (uint32_t)__popcll(b1 & lt) +
           4u * (uint32_t)__popcll(b2 & lt);
}

// ---------------------------------------------------------------------------
// FP64 pose helpers (pose_utils.cuh:13-188), w first.
// ---------------------------------------------------------------------------
__device__ inline void dev_rotmat_to_quat(const double* R, double* q) {
    const double tr = R[0] + R[4] + R[8];
    if (tr > 0.0) {
        const double s = 0.5 / sqrt(tr + 1.0);
        q[0] = 0.25 / s;
        q[1] = (R[7] - R[5]) * s;
        q[2] = (R[2] - R[6]) * s;
        q[3] = (R[3] - R[1]) * s;
    } else if (R[0] > R[4] && R[0] > R[8]) {
        const double s = 2.0 * sqrt(1.0 + R[0] - R[4] - R[8]);
        q[0] = (R[7] - R[5]) / s;
        q[1] = 0.25 * s;
        q[2] = (R[1] + R[3]) / s;
        q[3] = (R[2] + R[6]) / s;
    } else if (R[4] > R[8]) {
        const double s = 2.0 * sqrt(1.0 + R[4] - R[0] - R[8]);
        q[0] = (R[2] - R[6]) / s;
        q[1] = (R[1] + R[3]) / s;
        q[2] = 0.25 * s;
        q[3] = (R[5] + R[7]) / s;
    } else {
        const double s = 2.0 * sqrt(1.0 + R[8] - R[0] - R[4]);
        q[0] = (R[3] - R[1]) / s;
        q[1] = (R[2] + R[6]) / s;
        q[2] = (R[5] + R[7]) / s;
        q[3] = 0.25 * s;
    }
    const double nrm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int i = 0; i < 4; ++i) q[i] /= nrm;
}

__device__ inline void dev_quat_to_rotmat(const double* q, double* R) {
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    R[0] = 1.0 - 2.0 * (y * y + z * z);
    R[1] = 2.0 * (x * y - z * w);
    R[2] = 2.0 * (x * z + y * w);
    R[3] = 2.0 * (x * y + z * w);
    R[4] = 1.0 - 2.0 * (x * x + z * z);
    R[5] = 2.0 * (y * z - x * w);
    R[6] = 2.0 * (x * z - y * w);
    R[7] = 2.0 * (y * z + x * w);
    R[8] = 1.0 - 2.0 * (x * x + y * y);
}

__device__ inline void dev_quat_mul(const double* a, const double* b, double* o) {
    const double r0 = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    const double r1 = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    const double r2 = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    const double r3 = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    o[0] = r0; o[1] = r1; o[2] = r2; o[3] = r3;
}

__device__ inline void dev_quat_rot(const double* q, const double* p, double* o) {
    const double a = -q[1] * p[0] - q[2] * p[1] - q[3] * p[2];
    const double b = q[0] * p[0] + q[2] * p[2] - q[3] * p[1];
    const double c = q[0] * p[1] + q[3] * p[0] - q[1] * p[2];
    const double d = q[0] * p[2] + q[1] * p[1] - q[2] * p[0];
    o[0] = -a * q[1] + b * q[0] - c * q[3] + d * q[2];
    o[1] = -a * q[2] + c * q[0] - d * q[1] + b * q[3];
    o[2] = -a * q[3] + d * q[0] - b * q[2] + c * q[1];
}

__device__ inline void dev_quat_exp(const double* w, double* q) {
    const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double th = sqrt(th2);
    double re, im;
    if (th > 1e-6) {
        re = cos(0.5 * th);
        im = sin(0.5 * th) / th;
    } else {
        const double th4 = th2 * th2;
        re = 1.0 - th2 / 8.0 + th4 / 384.0;
        im = 0.5 - th2 / 48.0 + th4 / 3840.0;
        const double s = sqrt(re * re + im * im * th2);
        re /= s;
        im /= s;
    }
    q[0] = re;
    q[1] = im * w[0];
    q[2] = im * w[1];
    q[3] = im * w[2];
}

// ---------------------------------------------------------------------------
// prepare: the launch's first kernel.  It resets the per-frame counters and
// the header's (a kernel, so graphs hold no memset node; the grid has at
// least one thread per frame), and writes the normalised coordinates (FP64,
// exactly (u - cx) / fx) and the FP32 SoA arrays the scoring kernel streams,
// whose image coordinates are further divided by the threshold; padding
// gets u = NaN so it scores 0.  Nothing here reads what the reset writes.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kPtBlock) void k_prepare(const BatchDev* __restrict__ bp) {
    const BatchDev b = *bp;
    const int nf = b.hdr->n_frames;
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        FrameState* s = b.frames[f].state;
        s->n_cand = 0;
        s->topk_count = 0;
        s->best_score = DBL_MAX;
        s->best_id = -1;
        s->best_inliers = 0;
        s->total_candidates = 0;
        s->samples_drawn = 0;
        s->n_inliers = 0;
        s->status = TR_OK;
        s->active = 1;
        for (int i = 0; i < 12; ++i) s->best_pose[i] = (i % 5 == 0) ? 1.0 : 0.0;
    }
    if (blockIdx.x == 0) {
        if (threadIdx.x < kScoreQueues) b.hdr->q_head[threadIdx.x * kQHeadStride] = 0;
        if (threadIdx.x == 0) {
            b.hdr->n_active[0] = (uint32_t)nf;
            b.hdr->n_active[1] = 0;
            b.hdr->n_lm_live = 0;
        }
    }
    const int nb = b.hdr->n_pt_blocks;
    for (int it = blockIdx.x; it < nb; it += gridDim.x) {
        const BlockItem bi = b.pt_blocks[it];
        const FrameDesc& F = b.frames[bi.frame];
        const int i = bi.start + threadIdx.x;
        const int n = F.n, np = F.n_pad;
        float* p = F.pts32;
        if (i < n) {
            const double x = (F.uv[2 * i] - F.cx) / F.fx;
            const double y = (F.uv[2 * i + 1] - F.cy) / F.fy;
            F.xn[2 * i] = x;
            F.xn[2 * i + 1] = y;
            p[i] = (float)(x * F.inv_tau);
            p[np + i] = (float)(y * F.inv_tau);
            p[2 * np + i] = (float)F.xyz[3 * i];
            p[3 * np + i] = (float)F.xyz[3 * i + 1];
            p[4 * np + i] = (float)F.xyz[3 * i + 2];
        } else if (i < np) {
            p[i] = __builtin_nanf("");
            p[np + i] = 0.0f;
            p[2 * np + i] = 0.0f;
            p[3 * np + i] = 0.0f;
            p[4 * np + i] = 0.0f;
        }
    }
}

// The triple of global sample g under the frame's PROSAC schedule (Chum and
// Matas 2005): t = g + 1; past max_prosac uniform; else n = 3 + (events
// E_k <= t), and the sample is 3 of the top n when T' = E_k < t, else the
// n-th correspondence and 2 of the top n - 1.
__device__ void prosac_triple(const FrameDesc& F, uint64_t g, int32_t idx[3]) {
    const uint64_t t = g + 1;
    if (t > F.max_prosac) {
        sample_triple(F.seed, (uint32_t)F.n, g, idx);
        return;
    }
    int lo = 0, hi = F.prosac_len;  // first k with E_k > t, among the events
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (F.prosac[mid] <= t)
            lo = mid + 1;
        else
            hi = mid;
    }
    const uint32_t n = 3u + (uint32_t)lo;
    if (F.prosac[lo] < t) {
        sample_triple(F.seed, n, g, idx);
    } else {
        int32_t p[2];
        sample_pair(F.seed, n - 1u, g, p);
        idx[0] = p[0];
        idx[1] = p[1];
        idx[2] = (int32_t)(n - 1u);
    }
}

// ---------------------------------------------------------------------------
// sample + P3P: one minimal sample per lane, FP64.  Solutions go to the
// sample's four slots; the lanes of a wave then append their filled slots
// to the frame's candidate list with one atomic.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSampleBlock) void k_sample_p3p(const BatchDev* __restrict__ bp,
                                                             int round) {
    const BatchDev b = *bp;
    // The count of the next round starts here: k_choose of this round adds to it.
    if (blockIdx.x == 0 && threadIdx.x == 0) b.hdr->n_active[(round + 1) & 1] = 0;
    if (b.hdr->n_active[round & 1] == 0) return;
    const int nb = b.hdr->n_sample_blocks;
    for (int it = blockIdx.x; it < nb; it += gridDim.x) {
        const BlockItem bi = b.sample_blocks[it];
        const FrameDesc& F = b.frames[bi.frame];
        if (!F.state->active) continue;  // uniform over the block
        const int s = bi.start + threadIdx.x;
        // Samples of earlier rounds: this round continues the frame's stream.
        const uint64_t g = F.state->samples_drawn + (uint64_t)s;
        int ns = 0;
        int32_t idx[3] = {0, 0, 0};
        P3PSolutions P;
        if (s < F.H) {
            if (F.progressive)
                prosac_triple(F, g, idx);
            else
                sample_triple(F.seed, (uint32_t)F.n, g, idx);
            D3 xb[3], Xb[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double u = F.xn[2 * idx[k]], v = F.xn[2 * idx[k] + 1];
                const double nrm = sqrt(u * u + v * v + 1.0);
                xb[k] = {u / nrm, v / nrm, 1.0 / nrm};
                Xb[k] = {F.xyz[3 * idx[k]], F.xyz[3 * idx[k] + 1], F.xyz[3 * idx[k] + 2]};
            }
            ns = p3p_solve(xb[0], xb[1], xb[2], Xb[0], Xb[1], Xb[2], P);
        }
        uint32_t total;
        const uint32_t pre = wave_prefix_small((uint32_t)ns, &total);
        uint32_t base = 0;
        if (total) {
            if ((threadIdx.x & 63) == 0) base = atomicAdd(&F.state->n_cand, total);
            base = __shfl(base, 0, 64);
        }
        if (ns) {
            // The FP64 poses by slot 4 s + k (the re-rank, the mask and the
            // LM read them by slot), and the FP32 scoring copies, with rows
            // 0 and 1 (x and y) in units of the threshold like the scoring
            // arrays' image coordinates, at their candidate indices: the
            // 64 candidates of a scoring item are 3 KB in one piece.  Both
            // are written from registers.
            p3p_emit_all(P, F.slot64 + 48 * (size_t)s, F.cand32 + 12 * (size_t)(base + pre),
                         F.inv_tau);
            for (int k = 0; k < ns; ++k)
                F.cand_slot[base + pre + (uint32_t)k] = 4u * (uint32_t)s + (uint32_t)k;
        }
        if (s < F.H && F.dbg_samples && (int64_t)g < F.dbg_max_samples) {
            const double* out = F.slot64 + 48 * (size_t)s;
            F.dbg_samples[3 * g] = idx[0];
            F.dbg_samples[3 * g + 1] = idx[1];
            F.dbg_samples[3 * g + 2] = idx[2];
            F.dbg_nsol[g] = ns;
            for (int j = 0; j < 48; ++j)
                F.dbg_poses[48 * g + j] = (j < 12 * ns) ? out[j] : __builtin_nan("");
            for (int k = 0; k < kMaxSolutions; ++k) F.dbg_scores32[4 * g + k] = __builtin_nanf("");
        }
    }
}

// ---------------------------------------------------------------------------
// MSAC scoring, FP32.  A workgroup takes (64 W J) candidates against a range
// of points: each lane holds J candidate poses in registers, the points are
// staged through LDS in double-buffered SoA tiles of T points and read by
// every lane at the same address (broadcast).  Per (candidate, point) the
// lane accumulates w = clamp(1 - r^2 / tau^2, 0, 1), which is 0 for
// outliers; the candidate's score is tau^2 (N - sum w), the reference's
// sum over inliers of r^2 plus tau^2 per outlier.  Image coordinates and the
// pose's x and y rows are in units of tau, so w = 1 - r0^2 - r1^2, and the
// clamp is the output modifier of its last FMA (DX10 clamp mode maps NaN to
// 0).  Cheirality (z > 0) costs nothing: 1/z comes from rsq(z)^2, NaN for
// z < 0, which the clamp turns into 0.
// Each candidate's sum runs over its points in a fixed order, so the score
// does not depend on which workgroup or batch computed it.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float msac_weight(float z0, float z1, float z2, float u, float v) {
    const float rs = __builtin_amdgcn_rsqf(z2);
    const float iz = rs * rs;
    const float r0 = fmaf(z0, iz, -u);
    const float r1 = fmaf(z1, iz, -v);
    return __builtin_amdgcn_fmed3f(fmaf(-r0, r0, fmaf(-r1, r1, 1.0f)), 0.0f, 1.0f);
}

__device__ __forceinline__ float msac_term(const float* P, float u, float v, float X, float Y,
                                           float Z) {
    const float z2 = fmaf(P[8], X, fmaf(P[9], Y, fmaf(P[10], Z, P[11])));
    const float z0 = fmaf(P[0], X, fmaf(P[1], Y, fmaf(P[2], Z, P[3])));
    const float z1 = fmaf(P[4], X, fmaf(P[5], Y, fmaf(P[6], Z, P[7])));
    return msac_weight(z0, z1, z2, u, v);
}

// Next MSAC work item of a workgroup: from queue q (the workgroup's XCD,
// blockIdx % 8 under round-robin dispatch) until it runs dry, then from
// the other queues in turn.  [qb, qe) are queue q's bounds, kept in
// registers.  Returns -1 once every queue is exhausted; `left` bounds the
// probes, so every workgroup reaches the exit.
struct ScoreQueue {
    int q, left, qb, qe;
    __device__ void init(const BatchHeader* h) {
        q = blockIdx.x % kScoreQueues;
        left = kScoreQueues;
        qb = h->q_start[q];
        qe = h->q_start[q + 1];
    }
    __device__ int next(BatchHeader* h) {
        while (left > 0) {
            const int idx = qb + (int)atomicAdd(&h->q_head[q * kQHeadStride], 1u);
            if (idx < qe) return idx;
            q = (q + 1) % kScoreQueues;
            if (--left > 0) {
                qb = h->q_start[q];
                qe = h->q_start[q + 1];
            }
        }
        return -1;
    }
};

template <int J, int T, int W>
__global__ __launch_bounds__(64 * W) void k_msac(const BatchDev* __restrict__ bp, int round) {
    const BatchDev b = *bp;
    if (b.hdr->n_active[round & 1] == 0) return;
    constexpr int NT = 64 * W;            // threads
    constexpr int V4 = T / 4;             // float4 per array per tile
    constexpr int LPT = (V4 + NT - 1) / NT;
    static_assert(V4 % NT == 0 || NT % V4 == 0, "tile/threads mismatch");
    __shared__ float4 tile[2][5][V4];
    // The item index lives in buffer 1, which no wave touches between the
    // last barrier of an item and the first tile store of the next.
    int* s_item = reinterpret_cast<int*>(&tile[1][0][0]);

    const int tid = threadIdx.x;
    ScoreQueue queue;  // used by thread 0
    if (threadIdx.x == 0) queue.init(b.hdr);
    for (;;) {
        if (tid == 0) *s_item = queue.next(b.hdr);
        __syncthreads();
        // Uniform over the workgroup: item bounds and loop counts stay scalar.
        const int item = __builtin_amdgcn_readfirstlane(*s_item);
        __syncthreads();
        if (item < 0) break;
        const ScoreItem si = b.score_items[item];
        const FrameDesc& F = b.frames[si.frame];
        const int ncand = F.state->active ? (int)F.state->n_cand : 0;
        if (si.cbase >= ncand) continue;

        float P[J][12];
        float acc[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            // Lanes past the last candidate score a valid pose and drop it.
            const int c = min(si.cbase + tid + j * NT, ncand - 1);
            const float4* src = reinterpret_cast<const float4*>(F.cand32 + 12 * (size_t)c);
            const float4 a = src[0], bq = src[1], cq = src[2];
            P[j][0] = a.x; P[j][1] = a.y; P[j][2] = a.z; P[j][3] = a.w;
            P[j][4] = bq.x; P[j][5] = bq.y; P[j][6] = bq.z; P[j][7] = bq.w;
            P[j][8] = cq.x; P[j][9] = cq.y; P[j][10] = cq.z; P[j][11] = cq.w;
            acc[j] = 0.0f;
        }
        const int np = F.n_pad;
        const float4* gu = reinterpret_cast<const float4*>(F.pts32 + si.p0);
        const float4* gv = reinterpret_cast<const float4*>(F.pts32 + (size_t)np + si.p0);
        const float4* gx = reinterpret_cast<const float4*>(F.pts32 + 2 * (size_t)np + si.p0);
        const float4* gy = reinterpret_cast<const float4*>(F.pts32 + 3 * (size_t)np + si.p0);
        const float4* gz = reinterpret_cast<const float4*>(F.pts32 + 4 * (size_t)np + si.p0);
        const int npts = __builtin_amdgcn_readfirstlane(si.p1 - si.p0);  // multiple of kPad
        const int ntiles = (npts + T - 1) / T;
        float4 ru[LPT], rv[LPT], rx[LPT], ry[LPT], rz[LPT];

#define TR_TILE_LOAD(t_)                                                   \
    {                                                                      \
        const int v0_ = (t_) * V4;                                         \
        const int vend_ = min(V4, (npts - (t_) * T) / 4);                  \
        _Pragma("unroll") for (int l = 0; l < LPT; ++l) {                  \
            const int v = tid + l * NT;                                    \
            if (v < vend_) {                                               \
                ru[l] = gu[v0_ + v];                                       \
                rv[l] = gv[v0_ + v];                                       \
                rx[l] = gx[v0_ + v];                                       \
                ry[l] = gy[v0_ + v];                                       \
                rz[l] = gz[v0_ + v];                                       \
            }                                                              \
        }                                                                  \
    }
#define TR_TILE_STORE(buf_)                                                \
    {                                                                      \
        _Pragma("unroll") for (int l = 0; l < LPT; ++l) {                  \
            const int v = tid + l * NT;                                    \
            if (v < V4) {                                                  \
                tile[buf_][0][v] = ru[l];                                  \
                tile[buf_][1][v] = rv[l];                                  \
                tile[buf_][2][v] = rx[l];                                  \
                tile[buf_][3][v] = ry[l];                                  \
                tile[buf_][4][v] = rz[l];                                  \
            }                                                              \
        }                                                                  \
    }

        TR_TILE_LOAD(0)
        TR_TILE_STORE(0)
        __syncthreads();
        for (int t = 0; t < ntiles; ++t) {
            const int buf = t & 1;
            if (t + 1 < ntiles) TR_TILE_LOAD(t + 1)
            const int m4 = min(V4, (npts - t * T) / 4);
            const float4* tu = tile[buf][0];
            const float4* tv = tile[buf][1];
            const float4* tx = tile[buf][2];
            const float4* ty = tile[buf][3];
            const float4* tz = tile[buf][4];
#pragma unroll 2
            for (int q = 0; q < m4; ++q) {
                const float4 u = tu[q], v = tv[q], x = tx[q], y = ty[q], z = tz[q];
#pragma unroll
                for (int j = 0; j < J; ++j) {
                    acc[j] += msac_term(P[j], u.x, v.x, x.x, y.x, z.x);
                    acc[j] += msac_term(P[j], u.y, v.y, x.y, y.y, z.y);
                    acc[j] += msac_term(P[j], u.z, v.z, x.z, y.z, z.z);
                    acc[j] += msac_term(P[j], u.w, v.w, x.w, y.w, z.w);
                }
            }
            if (t + 1 < ntiles) TR_TILE_STORE(buf ^ 1)
            __syncthreads();
        }
#undef TR_TILE_LOAD
#undef TR_TILE_STORE
        float* out = F.partial + (size_t)si.chunk * F.cand_cap;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int c = si.cbase + tid + j * NT;
            if (c < ncand) out[c] = acc[j];
        }
    }
}
// ---------------------------------------------------------------------------
// MSAC scoring with the projection on the matrix cores.  z = [R|t] [X;1] is
// a K = 4 contraction, so v_mfma_f32_16x16x4f32 computes one component of
// z for 16 hypotheses x 16 points per instruction: A = the 16 hypotheses'
// pose rows (lane i + 16k holds row i, column k), B = the 16 points' (X, Y,
// Z, 1) (lane j + 16k holds component k of point j), and lane j + 16q
// receives z for hypotheses 4q..4q+3 at point j.  Three instructions give
// z0, z1, z2; the vector ALUs only evaluate the residual tail (6 operations
// and one rsq instead of 15 and one).  A wave holds M tiles of 16
// hypotheses; the points come from the same double-buffered LDS tiles as in
// k_msac.  Each (hypothesis, point column) partial runs over the points in
// order and the 16 columns are folded by a fixed butterfly, so scores are
// deterministic.
// ---------------------------------------------------------------------------
typedef float tr_v4f __attribute__((ext_vector_type(4)));

template <int M, int T, int W>
__global__ __launch_bounds__(64 * W) void k_msac_mfma(const BatchDev* __restrict__ bp,
                                                     int round) {
    const BatchDev b = *bp;
    if (b.hdr->n_active[round & 1] == 0) return;
    constexpr int NT = 64 * W;
    constexpr int V4 = T / 4;
    constexpr int LPT = (V4 + NT - 1) / NT;
    static_assert(V4 % NT == 0 || NT % V4 == 0, "tile/threads mismatch");
    __shared__ float4 tile[2][5][V4];
    int* s_item = reinterpret_cast<int*>(&tile[1][0][0]);

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int col = lane & 15;  // A row / B and D column
    const int kq = lane >> 4;   // A and B k index, D row group
    ScoreQueue queue;  // used by thread 0
    if (threadIdx.x == 0) queue.init(b.hdr);
    const tr_v4f zero = {0.0f, 0.0f, 0.0f, 0.0f};
    for (;;) {
        if (tid == 0) *s_item = queue.next(b.hdr);
        __syncthreads();
        // Uniform over the workgroup: item bounds and loop counts stay scalar.
        const int item = __builtin_amdgcn_readfirstlane(*s_item);
        __syncthreads();
        if (item < 0) break;
        const ScoreItem si = b.score_items[item];
        const FrameDesc& F = b.frames[si.frame];
        const int ncand = F.state->active ? (int)F.state->n_cand : 0;
        if (si.cbase >= ncand) continue;

        float a0[M], a1[M], a2[M];
        tr_v4f acc[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            // Rows past the last candidate score a valid pose and are dropped.
            const int c = min(si.cbase + (wave * M + m) * 16 + col, ncand - 1);
            const float* P = F.cand32 + 12 * (size_t)c;
            a0[m] = P[kq];
            a1[m] = P[4 + kq];
            a2[m] = P[8 + kq];
            acc[m] = zero;
        }
        const int np = F.n_pad;
        const float4* gu = reinterpret_cast<const float4*>(F.pts32 + si.p0);
        const float4* gv = reinterpret_cast<const float4*>(F.pts32 + (size_t)np + si.p0);
        const float4* gx = reinterpret_cast<const float4*>(F.pts32 + 2 * (size_t)np + si.p0);
        const float4* gy = reinterpret_cast<const float4*>(F.pts32 + 3 * (size_t)np + si.p0);
        const float4* gz = reinterpret_cast<const float4*>(F.pts32 + 4 * (size_t)np + si.p0);
        const int npts = __builtin_amdgcn_readfirstlane(si.p1 - si.p0);  // multiple of kPad
        const int ntiles = (npts + T - 1) / T;
        float4 ru[LPT], rv[LPT], rx[LPT], ry[LPT], rz[LPT];

#define TR_TILE_LOAD(t_)                                                   \
    {                                                                      \
        const int v0_ = (t_) * V4;                                         \
        const int vend_ = min(V4, (npts - (t_) * T) / 4);                  \
        _Pragma("unroll") for (int l = 0; l < LPT; ++l) {                  \
            const int v = tid + l * NT;                                    \
            if (v < vend_) {                                               \
                ru[l] = gu[v0_ + v];                                       \
                rv[l] = gv[v0_ + v];                                       \
                rx[l] = gx[v0_ + v];                                       \
                ry[l] = gy[v0_ + v];                                       \
                rz[l] = gz[v0_ + v];                                       \
            }                                                              \
        }                                                                  \
    }
#define TR_TILE_STORE(buf_)                                                \
    {                                                                      \
        _Pragma("unroll") for (int l = 0; l < LPT; ++l) {                  \
            const int v = tid + l * NT;                                    \
            if (v < V4) {                                                  \
                tile[buf_][0][v] = ru[l];                                  \
                tile[buf_][1][v] = rv[l];                                  \
                tile[buf_][2][v] = rx[l];                                  \
                tile[buf_][3][v] = ry[l];                                  \
                tile[buf_][4][v] = rz[l];                                  \
            }                                                              \
        }                                                                  \
    }

        TR_TILE_LOAD(0)
        TR_TILE_STORE(0)
        __syncthreads();
        for (int t = 0; t < ntiles; ++t) {
            const int buf = t & 1;
            if (t + 1 < ntiles) TR_TILE_LOAD(t + 1)
            const int mp = min(T, npts - t * T);
            const float* fu = reinterpret_cast<const float*>(tile[buf][0]);
            const float* fv = reinterpret_cast<const float*>(tile[buf][1]);
            // B operand: X, Y or Z of point `col` for k < 3, the constant 1 for k = 3.
            const float* fb = reinterpret_cast<const float*>(tile[buf][2 + (kq < 3 ? kq : 0)]);
            for (int p = 0; p < mp; p += 16) {
                const float bk = (kq < 3) ? fb[p + col] : 1.0f;
                const float u = fu[p + col], v = fv[p + col];
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    const tr_v4f z0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[m], bk, zero, 0, 0, 0);
                    const tr_v4f z1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[m], bk, zero, 0, 0, 0);
                    const tr_v4f z2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a2[m], bk, zero, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc[m][r] += msac_weight(z0[r], z1[r], z2[r], u, v);
                }
            }
            if (t + 1 < ntiles) TR_TILE_STORE(buf ^ 1)
            __syncthreads();
        }
#undef TR_TILE_LOAD
#undef TR_TILE_STORE
        // Fold the 16 point columns (lanes with the same kq), then lane
        // 16q writes hypotheses 4q..4q+3 of each tile.
        float* out = F.partial + (size_t)si.chunk * F.cand_cap;
#pragma unroll
        for (int m = 0; m < M; ++m) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float s = acc[m][r];
                s += __shfl_xor(s, 1, 64);
                s += __shfl_xor(s, 2, 64);
                s += __shfl_xor(s, 4, 64);
                s += __shfl_xor(s, 8, 64);
                const int c = si.cbase + (wave * M + m) * 16 + 4 * kq + r;
                if (col == 0 && c < ncand) out[c] = s;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// MSAC scoring with the points on the lanes and the division gated by an
// exact inlier test.
//
// Costs measured on gfx950 (wave64, clocks of the SIMD per instruction): an
// FP32 FMA/MUL/ADD whose operands are all VGPRs issues in about 2.3-2.8
// (the wave's two halves at once); one that reads an SGPR takes about 4.2;
// a transcendental (v_rcp, v_rsq) about 10.5; a branch that waits on the
// compare just before it costs about 10.  FP32 MFMA and VALU work do not
// overlap on a SIMD.  The kernel is shaped by those numbers:
//
// * For most (hypothesis, point) pairs the point is an outlier, whose weight
//   is 0 whatever the division gives.  With e = z_xy - uv z2 (image
//   coordinates in units of tau) the point is an inlier iff
//   |e|^2 < z2 |z2| (which also requires z2 > 0); only then is the weight
//   1 - |e|^2 / z2^2 needed.  The test is 15 all-VGPR VALU instructions.
// * A wavefront holds 64 J points (J per lane) in registers and takes the
//   item's hypotheses one at a time, the pose broadcast from LDS into VGPRs
//   (the item's poses are staged in the wave's LDS rows first).  For a
//   wrong hypothesis, 64 points rarely hold an inlier.
// * The rows of 64 points are tested in groups of kGatedRows: the group's
//   inlier ballots are ORed and one branch skips the whole group when there
//   is none; otherwise each row with an inlier (its ballot, kept in SGPRs)
//   computes its weights.
// * Each wave takes its own items from the queues: no barrier.
//
// Determinism: per hypothesis and 64 J points, each lane sums its J points
// in order, each row of 16 lanes is folded by a fixed DPP tree, each row's
// sum is added to the hypothesis's LDS slot for that row by one lane, group
// after group in order, and the four row slots are added in a fixed order.
// ---------------------------------------------------------------------------

// Sum over each row of 16 lanes, by a fixed tree (pairs, quads, half rows,
// rows); every lane of a row ends with its row's sum.
__device__ __forceinline__ float row_sum_f32(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, false));
    return v;
}

// A pointer every lane holds, as a wave-uniform (SGPR) value.
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

typedef const __attribute__((address_space(1))) float gfloat_t;

constexpr int kGatedRows = 4;  // rows of 64 points per inlier-test branch

template <int J, int CB, int W>
__global__ __launch_bounds__(64 * W) void k_msac_gated(const BatchDev* __restrict__ bp,
                                                      int round) {
    constexpr int G = kGatedRows;
    static_assert(J % G == 0 && J % 4 == 0, "items hold multiples of 256 points");
    static_assert(CB <= 64, "one hypothesis sum per lane");
    __shared__ float4 s_pose[W][3 * CB];
    // Per hypothesis of the item, the sums of the four 16-lane rows,
    // accumulated over the groups of points in order (one lane per slot).
    __shared__ float4 s_sum[W][CB];
    const BatchDev b = *bp;
    if (b.hdr->n_active[round & 1] == 0) return;
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    ScoreQueue queue;  // lane 0 of each wave takes the wave's items
    queue.init(b.hdr);
    for (;;) {
        int item = 0;
        if (lane == 0) item = queue.next(b.hdr);
        item = __builtin_amdgcn_readlane(item, 0);
        if (item < 0) break;
        const ScoreItem si = b.score_items[item];
        const FrameDesc& F = b.frames[si.frame];
        const int ncand =
            __builtin_amdgcn_readfirstlane(F.state->active ? (int)F.state->n_cand : 0);
        const int cbase = __builtin_amdgcn_readfirstlane(si.cbase);
        if (cbase >= ncand) continue;
        const int cend = min(cbase + CB, ncand);
        const int p0 = __builtin_amdgcn_readfirstlane(si.p0);
        const int p1 = __builtin_amdgcn_readfirstlane(si.p1);
        const int np = __builtin_amdgcn_readfirstlane(F.n_pad);
        gfloat_t* gu = (gfloat_t*)uniform_ptr(F.pts32);
        gfloat_t* gv = gu + (size_t)np;
        gfloat_t* gx = gu + 2 * (size_t)np;
        gfloat_t* gy = gu + 3 * (size_t)np;
        gfloat_t* gz = gu + 4 * (size_t)np;
        // The item's poses into this wave's LDS rows (lane j: pose j).  The
        // wave reads them back itself: its LDS operations complete in order.
        if (cbase + lane < cend) {
            const float4* src =
                reinterpret_cast<const float4*>(F.cand32 + 12 * (size_t)(cbase + lane));
            s_pose[wv][3 * lane] = src[0];
            s_pose[wv][3 * lane + 1] = src[1];
            s_pose[wv][3 * lane + 2] = src[2];
        }
        if (lane < CB) s_sum[wv][lane] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        __builtin_amdgcn_wave_barrier();
        for (int g0 = p0; g0 < p1; g0 += 64 * J) {
            // The group's points, J per lane, image coordinates negated;
            // kmax 64-point rows hold points (a multiple of 4).  Rows past
            // kmax in a test group that straddles it get u = NaN, which
            // fails the inlier test.
            const int kmax = min(J, (p1 - g0) / 64);
            float nu[J], nv[J], px[J], py[J], pz[J];
#pragma unroll
            for (int k = 0; k < J; ++k) {
                const int i = min(g0 + 64 * k + lane, p1 - 1);
                nu[k] = k < kmax ? -gu[i] : __builtin_nanf("");
                nv[k] = -gv[i];
                px[k] = gx[i];
                py[k] = gy[i];
                pz[k] = gz[i];
                // Kept negated in registers, so that e = z + nu z2 is a
                // two-operand v_fmac, not an FMA with a negated operand.
                asm volatile("" : "+v"(nu[k]), "+v"(nv[k]));
            }
            for (int c = cbase; c < cend; ++c) {
                const float4 a = s_pose[wv][3 * (c - cbase)];
                const float4 bq = s_pose[wv][3 * (c - cbase) + 1];
                const float4 cq = s_pose[wv][3 * (c - cbase) + 2];
                const float p[12] = {a.x, a.y, a.z, a.w, bq.x, bq.y, bq.z, bq.w,
                                     cq.x, cq.y, cq.z, cq.w};
                float acc = 0.0f;
#pragma unroll
                for (int k0 = 0; k0 < J; k0 += G) {
                    if (k0 % 4 == 0 && k0 >= kmax) break;
                    float sv[G], qv[G];
                    uint64_t m[G];
                    uint64_t any = 0;
#pragma unroll
                    for (int r = 0; r < G; ++r) {
                        const int k = k0 + r;
                        const float z2 = fmaf(p[8], px[k], fmaf(p[9], py[k], fmaf(p[10], pz[k], p[11])));
                        float e0 = fmaf(p[0], px[k], fmaf(p[1], py[k], fmaf(p[2], pz[k], p[3])));
                        float e1 = fmaf(p[4], px[k], fmaf(p[5], py[k], fmaf(p[6], pz[k], p[7])));
                        e0 = fmaf(nu[k], z2, e0);
                        e1 = fmaf(nv[k], z2, e1);
                        sv[r] = fmaf(e1, e1, e0 * e0);
                        qv[r] = z2 * fabsf(z2);
                        m[r] = __ballot(sv[r] < qv[r]);
                        any |= m[r];
                    }
                    if (__builtin_expect(any != 0, 0)) {
                        // The empty asm statements keep these branches
                        // (a wave with no inlier skips the reciprocal)
                        // from being if-converted.
                        asm volatile("");
#pragma unroll
                        for (int r = 0; r < G; ++r) {
                            if (m[r] != 0) {
                                asm volatile("");
                                const float w = __builtin_amdgcn_fmed3f(
                                    fmaf(-sv[r], __builtin_amdgcn_rcpf(qv[r]), 1.0f), 0.0f, 1.0f);
                                acc += sv[r] < qv[r] ? w : 0.0f;
                            }
                        }
                    }
                }
                // The row sums go to the hypothesis's four LDS slots, one
                // lane per slot: LDS operations of a wave complete in
                // order, so each slot adds the groups in order.
                const float rs = row_sum_f32(acc);
                if ((lane & 15) == 15)
                    atomicAdd(reinterpret_cast<float*>(&s_sum[wv][c - cbase]) + (lane >> 4), rs);
            }
        }
        __builtin_amdgcn_wave_barrier();
        float* out = F.partial + (size_t)si.chunk * F.cand_cap;
        if (cbase + lane < cend) {
            const float4 t = s_sum[wv][lane];
            out[cbase + lane] = ((t.x + t.y) + t.z) + t.w;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// The gated kernel with two hypotheses per pass over the points a lane
// holds: each point's two inlier tests are independent chains the wave
// interleaves, and one branch covers both hypotheses' rows of a group.
// Each hypothesis adds its rows in the same order, and its row sums go to
// its LDS slots group after group, as in k_msac_gated: the scores are the
// same bits.  An odd last hypothesis of an item runs paired with itself,
// and only one copy of its sums is kept.
template <int J, int CB, int W>
__global__ __launch_bounds__(64 * W) void k_msac_gated2(const BatchDev* __restrict__ bp,
                                                       int round) {
    constexpr int G = kGatedRows;
    static_assert(J % G == 0 && J % 4 == 0, "items hold multiples of 256 points");
    static_assert(CB <= 64, "one hypothesis sum per lane");
    __shared__ float4 s_pose[W][3 * CB];
    __shared__ float4 s_sum[W][CB];
    const BatchDev b = *bp;
    if (b.hdr->n_active[round & 1] == 0) return;
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    ScoreQueue queue;
    queue.init(b.hdr);
    for (;;) {
        int item = 0;
        if (lane == 0) item = queue.next(b.hdr);
        item = __builtin_amdgcn_readlane(item, 0);
        if (item < 0) break;
        const ScoreItem si = b.score_items[item];
        const FrameDesc& F = b.frames[si.frame];
        const int ncand =
            __builtin_amdgcn_readfirstlane(F.state->active ? (int)F.state->n_cand : 0);
        const int cbase = __builtin_amdgcn_readfirstlane(si.cbase);
        if (cbase >= ncand) continue;
        const int cend = min(cbase + CB, ncand);
        const int p0 = __builtin_amdgcn_readfirstlane(si.p0);
        const int p1 = __builtin_amdgcn_readfirstlane(si.p1);
        const int np = __builtin_amdgcn_readfirstlane(F.n_pad);
        gfloat_t* gu = (gfloat_t*)uniform_ptr(F.pts32);
        gfloat_t* gv = gu + (size_t)np;
        gfloat_t* gx = gu + 2 * (size_t)np;
        gfloat_t* gy = gu + 3 * (size_t)np;
        gfloat_t* gz = gu + 4 * (size_t)np;
        if (cbase + lane < cend) {
            const float4* src =
                reinterpret_cast<const float4*>(F.cand32 + 12 * (size_t)(cbase + lane));
            s_pose[wv][3 * lane] = src[0];
            s_pose[wv][3 * lane + 1] = src[1];
            s_pose[wv][3 * lane + 2] = src[2];
        }
        if (lane < CB) s_sum[wv][lane] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        __builtin_amdgcn_wave_barrier();
        for (int g0 = p0; g0 < p1; g0 += 64 * J) {
            const int kmax = min(J, (p1 - g0) / 64);
            float nu[J], nv[J], px[J], py[J], pz[J];
#pragma unroll
            for (int k = 0; k < J; ++k) {
                const int i = min(g0 + 64 * k + lane, p1 - 1);
                nu[k] = k < kmax ? -gu[i] : __builtin_nanf("");
                nv[k] = -gv[i];
                px[k] = gx[i];
                py[k] = gy[i];
                pz[k] = gz[i];
                asm volatile("" : "+v"(nu[k]), "+v"(nv[k]));
            }
            for (int c = cbase; c < cend; c += 2) {
                const int c1 = min(c + 1, cend - 1);
                float p[12], p1v[12];
                {
                    const float4 a = s_pose[wv][3 * (c - cbase)];
                    const float4 bq = s_pose[wv][3 * (c - cbase) + 1];
                    const float4 cq = s_pose[wv][3 * (c - cbase) + 2];
                    const float4 a1 = s_pose[wv][3 * (c1 - cbase)];
                    const float4 b1 = s_pose[wv][3 * (c1 - cbase) + 1];
                    const float4 c1q = s_pose[wv][3 * (c1 - cbase) + 2];
                    const float t0[12] = {a.x, a.y, a.z, a.w, bq.x, bq.y, bq.z, bq.w,
                                          cq.x, cq.y, cq.z, cq.w};
                    const float t1[12] = {a1.x, a1.y, a1.z, a1.w, b1.x, b1.y, b1.z, b1.w,
                                          c1q.x, c1q.y, c1q.z, c1q.w};
#pragma unroll
                    for (int e = 0; e < 12; ++e) {
                        p[e] = t0[e];
                        p1v[e] = t1[e];
                    }
                }
                float acc0 = 0.0f, acc1 = 0.0f;
#pragma unroll
                for (int k0 = 0; k0 < J; k0 += G) {
                    if (k0 % 4 == 0 && k0 >= kmax) break;
                    float sv0[G], qv0[G], sv1[G], qv1[G];
                    uint64_t m0[G], m1[G];
                    uint64_t any = 0;
#pragma unroll
                    for (int r = 0; r < G; ++r) {
                        const int k = k0 + r;
                        const float z2 = fmaf(p[8], px[k], fmaf(p[9], py[k], fmaf(p[10], pz[k], p[11])));
                        const float y2 = fmaf(p1v[8], px[k], fmaf(p1v[9], py[k], fmaf(p1v[10], pz[k], p1v[11])));
                        float e0 = fmaf(p[0], px[k], fmaf(p[1], py[k], fmaf(p[2], pz[k], p[3])));
                        float f0 = fmaf(p1v[0], px[k], fmaf(p1v[1], py[k], fmaf(p1v[2], pz[k], p1v[3])));
                        float e1 = fmaf(p[4], px[k], fmaf(p[5], py[k], fmaf(p[6], pz[k], p[7])));
                        float f1 = fmaf(p1v[4], px[k], fmaf(p1v[5], py[k], fmaf(p1v[6], pz[k], p1v[7])));
                        e0 = fmaf(nu[k], z2, e0);
                        f0 = fmaf(nu[k], y2, f0);
                        e1 = fmaf(nv[k], z2, e1);
                        f1 = fmaf(nv[k], y2, f1);
                        sv0[r] = fmaf(e1, e1, e0 * e0);
                        sv1[r] = fmaf(f1, f1, f0 * f0);
                        qv0[r] = z2 * fabsf(z2);
                        qv1[r] = y2 * fabsf(y2);
                        m0[r] = __ballot(sv0[r] < qv0[r]);
                        m1[r] = __ballot(sv1[r] < qv1[r]);
                        any |= m0[r] | m1[r];
                    }
                    if (__builtin_expect(any != 0, 0)) {
                        asm volatile("");
#pragma unroll
                        for (int r = 0; r < G; ++r) {
                            if (m0[r] != 0) {
                                asm volatile("");
                                const float w = __builtin_amdgcn_fmed3f(
                                    fmaf(-sv0[r], __builtin_amdgcn_rcpf(qv0[r]), 1.0f), 0.0f, 1.0f);
                                acc0 += sv0[r] < qv0[r] ? w : 0.0f;
                            }
                            if (m1[r] != 0) {
                                asm volatile("");
                                const float w = __builtin_amdgcn_fmed3f(
                                    fmaf(-sv1[r], __builtin_amdgcn_rcpf(qv1[r]), 1.0f), 0.0f, 1.0f);
                                acc1 += sv1[r] < qv1[r] ? w : 0.0f;
                            }
                        }
                    }
                }
                const float rs0 = row_sum_f32(acc0);
                const float rs1 = row_sum_f32(acc1);
                if ((lane & 15) == 15) {
                    atomicAdd(reinterpret_cast<float*>(&s_sum[wv][c - cbase]) + (lane >> 4), rs0);
                    if (c1 != c)
                        atomicAdd(reinterpret_cast<float*>(&s_sum[wv][c1 - cbase]) + (lane >> 4),
                                  rs1);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        float* out = F.partial + (size_t)si.chunk * F.cand_cap;
        if (cbase + lane < cend) {
            const float4 t = s_sum[wv][lane];
            out[cbase + lane] = ((t.x + t.y) + t.z) + t.w;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------------------
// select: per frame, the kTopK lowest FP32 scores.  Key = (score order
// bits << 32) | (sample << 2 | solution), so ties go to the earliest sample
// and the selection does not depend on the order candidates were appended.
// ---------------------------------------------------------------------------
__device__ inline float key_to_float(uint64_t key) {
    const uint32_t u = (uint32_t)(key >> 32);
    const uint32_t bits = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    return __uint_as_float(bits);
}

constexpr int kSelectThreads = 256;

// One thread per candidate: the chunk partials, in chunk order, into chunk 0.
__global__ __launch_bounds__(kPtBlock) void k_sumscores(const BatchDev* __restrict__ bp,
                                                      int round) {
    const BatchDev b = *bp;
    if (b.hdr->n_active[round & 1] == 0) return;
    const int nb = b.hdr->n_cand_blocks;
    for (int it = blockIdx.x; it < nb; it += gridDim.x) {
        const BlockItem bi = b.cand_blocks[it];
        const FrameDesc& F = b.frames[bi.frame];
        const int nch = F.n_chunks;
        const int c = bi.start + threadIdx.x;
        if (nch < 2 || !F.state->active || c >= (int)F.state->n_cand) continue;
        const size_t cap = (size_t)F.cand_cap;
        float s = F.partial[c];
        for (int ch = 1; ch < nch; ++ch) s += F.partial[ch * cap + c];
        F.partial[c] = s;
    }
}

__global__ __launch_bounds__(kSelectThreads) void k_select(const BatchDev* __restrict__ bp,
                                                           int round) {
    const BatchDev b = *bp;
    __shared__ uint64_t s_min[kSelectThreads / 64];
    if (b.hdr->n_active[round & 1] == 0) return;
    const int nf = b.hdr->n_frames;
    const int tid = threadIdx.x;
    for (int f = blockIdx.x; f < nf; f += gridDim.x) {
        const FrameDesc& F = b.frames[f];
        if (!F.state->active) continue;  // uniform over the block
        const int ncand = (int)F.state->n_cand;
        const uint64_t round_base = F.state->samples_drawn;
        const float base = (float)((double)F.n * F.tau2);
        const float tau2 = (float)F.tau2;
        uint64_t key[kTopK];
        uint32_t idx[kTopK];
#pragma unroll
        for (int k = 0; k < kTopK; ++k) {
            key[k] = ~0ull;
            idx[k] = 0;
        }
        for (int c = tid; c < ncand; c += kSelectThreads) {
            // Minus the candidate's inlier weight (all chunks, folded by
            // k_sumscores): ascending order is ascending score.
            const float s = -F.partial[c];
            const uint32_t slot = F.cand_slot[c];
            const uint32_t id = (uint32_t)(round_base << 2) + slot;
            if (F.dbg_scores32) {
                const uint64_t g = id >> 2;
                if ((int64_t)g < F.dbg_max_samples) F.dbg_scores32[id] = fmaf(tau2, s, base);
            }
            uint64_t x = ((uint64_t)float_order_key(s) << 32) | id;
            uint32_t xi = slot;
#pragma unroll
            for (int k = 0; k < kTopK; ++k) {
                const bool sw = x < key[k];
                const uint64_t lo = sw ? x : key[k], hi = sw ? key[k] : x;
                const uint32_t ilo = sw ? xi : idx[k], ihi = sw ? idx[k] : xi;
                key[k] = lo;
                idx[k] = ilo;
                x = hi;
                xi = ihi;
            }
        }
        // Merge the per-thread sorted lists: kTopK rounds of block min.
        int taken = 0;
        for (int r = 0; r < kTopK; ++r) {
            const uint64_t m = wave_min_u64(key[0]);
            if ((tid & 63) == 0) s_min[tid >> 6] = m;
            __syncthreads();
            uint64_t best = s_min[0];
            for (int w = 1; w < kSelectThreads / 64; ++w) best = s_min[w] < best ? s_min[w] : best;
            __syncthreads();
            if (best == ~0ull) break;
            if (key[0] == best) {
                FrameState* st = F.state;
                st->topk_idx[r] = idx[0];
                st->topk_id[r] = (uint32_t)best;
                st->topk_score32[r] = base + key_to_float(best);
#pragma unroll
                for (int k = 0; k + 1 < kTopK; ++k) {
                    key[k] = key[k + 1];
                    idx[k] = idx[k + 1];
                }
                key[kTopK - 1] = ~0ull;
            }
            ++taken;
        }
        if (tid == 0) F.state->topk_count = (uint32_t)taken;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// rerank: FP64 MSAC (the reference's formula) of the leaders over a range of
// points; partial sums per (chunk, leader).
// ---------------------------------------------------------------------------
constexpr int kRerankThreads = 256;

__global__ __launch_bounds__(kRerankThreads) void k_rerank(const BatchDev* __restrict__ bp,
                                                           int round) {
    const BatchDev b = *bp;
    __shared__ double s_pose[kRrSlots][12];
    __shared__ double s_acc[kRerankThreads / 64][kRrSlots];
    __shared__ uint32_t s_cnt[kRerankThreads / 64][kRrSlots];
    if (b.hdr->n_active[round & 1] == 0) return;
    const int ni = b.hdr->n_rr_items;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int it = blockIdx.x; it < ni; it += gridDim.x) {
        const RangeItem ri = b.rr_items[it];
        const FrameDesc& F = b.frames[ri.frame];
        const int K = F.state->active ? (int)F.state->topk_count : 0;
        // The initial pose takes slot kTopK in the first round.
        const int Kall = K + ((round == 0 && F.score_init) ? 1 : 0);
        if (Kall == 0) continue;  // uniform over the block
        // The poses, once per item, in LDS.
        if (tid < K * 12) {
            s_pose[tid / 12][tid % 12] = F.slot64[12 * (size_t)F.state->topk_idx[tid / 12] + tid % 12];
        } else if (tid == K * 12 && Kall > K) {
            double R[9];
            dev_quat_to_rotmat(F.init, R);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                s_pose[K][4 * r] = R[3 * r];
                s_pose[K][4 * r + 1] = R[3 * r + 1];
                s_pose[K][4 * r + 2] = R[3 * r + 2];
                s_pose[K][4 * r + 3] = F.init[4 + r];
            }
        }
        __syncthreads();
        const double tau2 = F.tau2;
        double acc[kRrSlots];
        uint32_t cnt[kRrSlots];
#pragma unroll
        for (int k = 0; k < kRrSlots; ++k) {
            acc[k] = 0.0;
            cnt[k] = 0;
        }
        // Every pose against each point (msac.cuh:18-56 in FP64).
        for (int i = ri.p0 + tid; i < ri.p1; i += kRerankThreads) {
            const double X0 = F.xyz[3 * i], X1 = F.xyz[3 * i + 1], X2 = F.xyz[3 * i + 2];
            const double x0 = F.xn[2 * i], x1 = F.xn[2 * i + 1];
#pragma unroll
            for (int k = 0; k < kRrSlots; ++k) {
                if (k >= Kall) break;
                const double* p = s_pose[k];
                const double z0 = p[0] * X0 + p[1] * X1 + p[2] * X2 + p[3];
                const double z1 = p[4] * X0 + p[5] * X1 + p[6] * X2 + p[7];
                const double z2 = p[8] * X0 + p[9] * X1 + p[10] * X2 + p[11];
                const double iz = 1.0 / z2;
                const double r0 = z0 * iz - x0;
                const double r1 = z1 * iz - x1;
                const double r2 = r0 * r0 + r1 * r1;
                if (r2 < tau2 && z2 > 0.0) {
                    acc[k] += r2;
                    ++cnt[k];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kRrSlots; ++k) {
            if (k >= Kall) break;
            const double a = wave_sum_f64(acc[k]);
            const uint32_t c = wave_sum_u32(cnt[k]);
            if (lane == 0) {
                s_acc[wave][k] = a;
                s_cnt[wave][k] = c;
            }
        }
        __syncthreads();
        if (tid < Kall) {
            double a = 0.0;
            uint32_t c = 0;
            for (int w = 0; w < kRerankThreads / 64; ++w) {
                a += s_acc[w][tid];
                c += s_cnt[w][tid];
            }
            const int slot = tid < K ? tid : kTopK;
            F.rr_acc[(size_t)ri.chunk * kRrSlots + slot] = a;
            F.rr_cnt[(size_t)ri.chunk * kRrSlots + slot] = c;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// choose: per frame, the FP64 best of the leaders, kept over rounds; then
// the RANSAC pose through its quaternion for the mask and the refinement.
// ---------------------------------------------------------------------------
// Samples the stop rule asks for at inlier ratio w (turbo_ransac.cu:223-231,
// with success_prob and dyn_num_trials_mult honoured): no inlier asks for
// the cap, w = 1 for none.  So does a w whose cube is below the FP64
// resolution of 1 - w^3 (log(1 - w^3) = 0 would ask for no sample at all).
__device__ uint64_t dev_needed_samples(double w, const FrameDesc& F) {
    if (w <= 0.0) return F.max_samples;
    if (w >= 1.0) return 0;
    const double miss = log(1.0 - w * w * w);
    if (!(miss < 0.0)) return F.max_samples;
    const double v = ceil(log(1.0 - F.success_prob) / miss * F.trials_mult);
    if (!(v < 1.8e19)) return F.max_samples;
    return (uint64_t)fmax(0.0, v);
}

// One wave per frame: the lanes fold the re-rank chunks of every slot (a
// fixed-order butterfly, the same sum on every lane), then lane 0 decides.
__global__ __launch_bounds__(64) void k_choose(const BatchDev* __restrict__ bp, int round) {
    const BatchDev b = *bp;
    if (b.hdr->n_active[round & 1] == 0) return;
    // The MSAC kernel of this round is done: its queues restart for the next.
    if (blockIdx.x == 0 && threadIdx.x < kScoreQueues)
        b.hdr->q_head[threadIdx.x * kQHeadStride] = 0;
    const int nf = b.hdr->n_frames;
    const int lane = threadIdx.x;
    for (int f = blockIdx.x; f < nf; f += gridDim.x) {
        const FrameDesc& F = b.frames[f];
        FrameState* st = F.state;
        if (!st->active) continue;  // uniform over the wave
        const int K = (int)st->topk_count;
        const bool init = round == 0 && F.score_init;
        double fa[kRrSlots];
        uint32_t fc[kRrSlots];
#pragma unroll
        for (int k = 0; k < kRrSlots; ++k) {
            fa[k] = 0.0;
            fc[k] = 0;
        }
        for (int ch = lane; ch < F.rr_chunks; ch += 64) {
#pragma unroll
            for (int k = 0; k < kRrSlots; ++k) {
                if (k < K || (k == kTopK && init)) {
                    fa[k] += F.rr_acc[(size_t)ch * kRrSlots + k];
                    fc[k] += F.rr_cnt[(size_t)ch * kRrSlots + k];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kRrSlots; ++k) {
            fa[k] = wave_sum_f64(fa[k]);
            fc[k] = wave_sum_u32(fc[k]);
        }
        if (lane != 0) continue;
        // The initial pose (score_initial_model): the earliest hypothesis,
        // id -2, scored in the first round before the leaders compete.
        if (init) {
            const double a = fa[kTopK];
            const uint32_t c = fc[kTopK];
            st->best_score = a + (double)(F.n - (int)c) * F.tau2;
            st->best_id = -2;
            st->best_inliers = c;
            double R[9];
            dev_quat_to_rotmat(F.init, R);
            for (int r = 0; r < 3; ++r) {
                for (int j = 0; j < 3; ++j) st->best_pose[4 * r + j] = R[3 * r + j];
                st->best_pose[4 * r + 3] = F.init[4 + r];
            }
        }
        double best = st->best_score;
        int64_t best_id = st->best_id;
        int kbest = -1;
        uint32_t kin = 0;
        for (int k = 0; k < K; ++k) {
            const double a = fa[k];
            const uint32_t c = fc[k];
            const double score = a + (double)(F.n - (int)c) * F.tau2;
            st->topk_score64[k] = score;
            st->topk_inliers[k] = c;
            const int64_t id = (int64_t)st->topk_id[k];
            if (score < best || (score == best && best_id >= 0 && id < best_id)) {
                best = score;
                best_id = id;
                kbest = k;
                kin = c;
            }
        }
        if (kbest >= 0) {
            st->best_score = best;
            st->best_id = best_id;
            st->best_inliers = kin;
            const double* P = F.slot64 + 12 * (size_t)st->topk_idx[kbest];
            for (int j = 0; j < 12; ++j) st->best_pose[j] = P[j];
        }
        st->total_candidates += st->n_cand;
        st->samples_drawn += (uint64_t)F.H;
        double q[4] = {1.0, 0.0, 0.0, 0.0};
        if (st->best_id != -1) {
            const double R[9] = {st->best_pose[0], st->best_pose[1], st->best_pose[2],
                                 st->best_pose[4], st->best_pose[5], st->best_pose[6],
                                 st->best_pose[8], st->best_pose[9], st->best_pose[10]};
            dev_rotmat_to_quat(R, q);
        }
        double Rq[9];
        dev_quat_to_rotmat(q, Rq);
        for (int i = 0; i < 4; ++i) st->mask_q[i] = q[i];
        for (int i = 0; i < 9; ++i) st->mask_R[i] = Rq[i];
        st->mask_t[0] = st->best_pose[3];
        st->mask_t[1] = st->best_pose[7];
        st->mask_t[2] = st->best_pose[11];
        // Stop rule (turbo_ransac.cu:223-257): another round unless the
        // samples drawn reach the cap or exceed what the best inlier ratio
        // asks for (and the minimum).
        const uint64_t drawn = st->samples_drawn;
        bool more = false;
        if (F.adaptive && drawn < F.max_samples) {
            const double w = st->best_id != -1 ? (double)st->best_inliers / (double)F.n : 0.0;
            more = !(drawn >= F.min_samples && drawn > dev_needed_samples(w, F));
        }
        if (more) {
            st->n_cand = 0;
            st->topk_count = 0;
            atomicAdd(&b.hdr->n_active[(round + 1) & 1], 1u);
        } else {
            st->active = 0;
        }
    }
}

// ---------------------------------------------------------------------------
// mask: r^2 < tau^2 and z > 0 for the RANSAC pose (FP64), and the inlier
// count of every block of kPtBlock points (the offsets of the compaction).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kPtBlock) void k_mask(const BatchDev* __restrict__ bp) {
    const BatchDev b = *bp;
    __shared__ uint32_t s_cnt[kPtBlock / 64];
    const int nb = b.hdr->n_pt_blocks;
    const int tid = threadIdx.x;
    for (int it = blockIdx.x; it < nb; it += gridDim.x) {
        const BlockItem bi = b.pt_blocks[it];
        const FrameDesc& F = b.frames[bi.frame];
        const int i = bi.start + tid;
        bool in = false;
        if (i < F.n) {
            const FrameState* st = F.state;
            const double* R = st->mask_R;
            const double* t = st->mask_t;
            const double X0 = F.xyz[3 * i], X1 = F.xyz[3 * i + 1], X2 = F.xyz[3 * i + 2];
            const double Z0 = R[0] * X0 + R[1] * X1 + R[2] * X2 + t[0];
            const double Z1 = R[3] * X0 + R[4] * X1 + R[5] * X2 + t[1];
            const double Z2 = R[6] * X0 + R[7] * X1 + R[8] * X2 + t[2];
            const double r0 = Z0 / Z2 - F.xn[2 * i];
            const double r1 = Z1 / Z2 - F.xn[2 * i + 1];
            in = r0 * r0 + r1 * r1 < F.tau2 && Z2 > 0.0;
            F.mask[i] = in ? 1 : 0;
        }
        const uint32_t c = (uint32_t)__popcll(__ballot(in));
        if ((tid & 63) == 0) s_cnt[tid >> 6] = c;
        __syncthreads();
        if (tid == 0 && bi.start < F.n) {
            uint32_t sum = 0;
#pragma unroll
            for (int w = 0; w < kPtBlock / 64; ++w) sum += s_cnt[w];
            F.blk_count[bi.start / kPtBlock] = sum;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// LM (entry.cu:87-132, lm.cuh:15-171, jaco.cuh).  k_lm_compact gathers each
// frame's inliers (pixels / f, world) in point order and starts the frame's
// LM state at the RANSAC pose.  k_lm_iter then takes one LM iteration per
// launch: every workgroup evaluates the robust cost and the normal
// equations of the proposed pose over one chunk of kLmChunk inliers, and
// the last workgroup of a frame to finish folds the chunk sums in chunk
// order and takes the step (accept or reject, damping, Cholesky solve, next
// proposal).  An accepted step reuses the pass's normal equations and a
// rejected one keeps the previous ones, exactly as the reference
// recomputes them only after an accepted step.  The final launch finishes
// every frame still running in one workgroup per frame, with the same
// chunk sums folded in the same order, so a frame's result does not depend
// on how many launches served it.
// ---------------------------------------------------------------------------
// Loss and IRLS weight by loss type, the type a template argument (the
// Cauchy terms of the inlier loop are evaluated there).
// TR-IRLS weight of Le and Zach ("Robust Fitting with Truncated Least
// Squares: A Bilevel Optimization Approach", 3DV 2021) as published in
// PoseLib's TruncatedLossLeZach, whose enum and comment the reference's
// options carry (utils.h:50-51): at x = r^2 / s^2 >= 1 the weight is
// (z* - zbar) / rho with z* = 1, zbar the minimiser over [0, 1] of the
// penalised quadratic (1 + mu rho) z^2 / 2 - (x + mu rho z* - rho x / 2) z,
// and rho = (2 (x - 1) + sqrt(4 (x - 1)^2 mu^2 + 2 mu (x - 1))) / mu, with
// mu = 1/2; inside the threshold the weight is 1/2.  At x = 1 the
// published form divides 0 by 0; its limit from above, 1/2, is taken.
__host__ __device__ inline double le_zach_weight(double x) {
    if (x <= 1.0) return 0.5;
    const double mu = 0.5;
    const double d = x - 1.0;
    const double rho = (2.0 * d + sqrt(4.0 * d * d * mu * mu + 2.0 * mu * d)) / mu;
    const double a = (x + mu * rho - 0.5 * rho * x) / (1.0 + mu * rho);
    const double zbar = fmax(0.0, fmin(a, 1.0));
    return (1.0 - zbar) / rho;
}

struct LmLoss {
    int type;
    double s, s2, inv_s2;
    template <int LT>
    __device__ double loss(double r2) const {
        if (LT == TR_LOSS_TRIVIAL) return r2;
        if (LT == TR_LOSS_TRUNCATED || LT == TR_LOSS_TRUNCATED_LE_ZACH) return fmin(r2, s2);
        if (LT == TR_LOSS_HUBER) {
            const double r = sqrt(r2);
            return (r <= s) ? r2 : s * (2.0 * r - s);
        }
        return log1p(r2 * inv_s2);
    }
    template <int LT>
    __device__ double weight(double r2) const {
        if (LT == TR_LOSS_TRIVIAL) return 1.0;
        if (LT == TR_LOSS_TRUNCATED) return (r2 < s2) ? 1.0 : 0.0;
        if (LT == TR_LOSS_TRUNCATED_LE_ZACH) return le_zach_weight(r2 * inv_s2);
        if (LT == TR_LOSS_HUBER) {
            const double r = sqrt(r2);
            return (r <= s) ? 1.0 : s / r;
        }
        return fmax(DBL_MIN, inv_s2 / (1.0 + r2 * inv_s2));
    }
};

struct Cam4 {
    double fx, fy, cx, cy;
};

// Sums 32 per-lane values over the wave with 32 shuffles (reduce-scatter):
// each step trades half of the values with the partner lane, so afterwards
// lane l holds the total of value index l >> 1.  Fixed order, deterministic.
__device__ __forceinline__ double wave_transpose_reduce(double (&v)[32]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int bit = 5, half = 16; bit >= 1; --bit, half >>= 1) {
        const bool hi = (lane >> bit) & 1;
#pragma unroll
        for (int i = 0; i < half; ++i) {
            const double a = v[i], c = v[half + i];
            const double send = hi ? a : c;
            const double keep = hi ? c : a;
            v[i] = keep + __shfl_xor(send, 1 << bit, 64);
        }
    }
    return v[0] + __shfl_xor(v[0], 1, 64);
}

// The inlier loop of one chunk for one loss type (a template argument, so
// the loop carries no switch).  The Cauchy cost of the unit-weight terms
// is the log of their product: sum log1p(x_i) = log prod (1 + x_i), kept
// as a mantissa in [0.5, 1) and a binary exponent, with one log per lane
// at the end instead of one log1p per term; terms of any other weight add
// wt * log1p(x) as before.  Unit weights and no weights thus take the same
// path and give the same bits.
template <int LT>
__device__ __forceinline__ void lm_chunk_loop(const double* __restrict__ lx,
                                              const double* __restrict__ lX,
                                              const double* __restrict__ lw, int p0, int p1,
                                              Cam4 K, LmLoss L, const double (&R)[9], double t0,
                                              double t1, double t2, double (&acc)[32]) {
#pragma clang fp contract(fast)
    double pm = 1.0;  // Cauchy: product of (1 + r2 / s^2), mantissa
    int pe = 0;       // and exponent
    // The next inlier's inputs are loaded while this one is evaluated, so
    // the loads' latency overlaps the arithmetic.
    int i = p0 + (int)(threadIdx.x % kLmThreads);
    double nX0 = 0.0, nX1 = 0.0, nX2 = 0.0, nx0 = 0.0, nx1 = 0.0, nwt = 1.0;
    if (i < p1) {
        nX0 = lX[3 * i]; nX1 = lX[3 * i + 1]; nX2 = lX[3 * i + 2];
        nx0 = lx[2 * i]; nx1 = lx[2 * i + 1];
        if (lw) nwt = lw[i];
    }
    for (; i < p1; i += kLmThreads) {
        const double X0 = nX0, X1 = nX1, X2 = nX2, u = nx0, v = nx1;
        // Term weights (jaco.cuh:63, :151): 1 without them.
        const double wt = nwt;
        const int in = i + kLmThreads;
        if (in < p1) {
            nX0 = lX[3 * in]; nX1 = lX[3 * in + 1]; nX2 = lX[3 * in + 2];
            nx0 = lx[2 * in]; nx1 = lx[2 * in + 1];
            if (lw) nwt = lw[in];
        }
        const double Z0 = R[0] * X0 + R[1] * X1 + R[2] * X2 + t0;
        const double Z1 = R[3] * X0 + R[4] * X1 + R[5] * X2 + t1;
        const double Z2 = R[6] * X0 + R[7] * X1 + R[8] * X2 + t2;
        if (!(Z2 >= 0.0)) continue;
        const double iz = 1.0 / Z2;
        const double z0 = Z0 * iz, z1 = Z1 * iz;
        const double r0 = K.fx * z0 + K.cx - u;
        const double r1 = K.fy * z1 + K.cy - v;
        const double r2 = r0 * r0 + r1 * r1;
        double w;
        if (LT == TR_LOSS_CAUCHY) {
            const double x = r2 * L.inv_s2;
            const double a = 1.0 + x;
            if (wt == 1.0) {
                int ex;
                pm = frexp(pm * a, &ex);
                pe += ex;
            } else {
                acc[0] += wt * log1p(x);
            }
            w = wt * fmax(DBL_MIN, L.inv_s2 / a);
        } else {
            acc[0] += wt * L.template loss<LT>(r2);
            w = wt * L.template weight<LT>(r2);
            if (w == 0.0) continue;
        }
        const double P0a = K.fx * iz, P0c = -K.fx * z0 * iz;
        const double P1b = K.fy * iz, P1c = -K.fy * z1 * iz;
        double D0[3], D1[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            D0[c] = P0a * R[c] + P0c * R[6 + c];
            D1[c] = P1b * R[3 + c] + P1c * R[6 + c];
        }
        const double J0[6] = {X1 * D0[2] - X2 * D0[1], X2 * D0[0] - X0 * D0[2],
                              X0 * D0[1] - X1 * D0[0], D0[0], D0[1], D0[2]};
        const double J1[6] = {X1 * D1[2] - X2 * D1[1], X2 * D1[0] - X0 * D1[2],
                              X0 * D1[1] - X1 * D1[0], D1[0], D1[1], D1[2]};
        // Two FMAs per sum: acc + w J0a J0c + w J1a J1c.
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            const double wa0 = w * J0[a], wa1 = w * J1[a];
#pragma unroll
            for (int c = 0; c <= a; ++c) {
                double& s = acc[1 + a * (a + 1) / 2 + c];
                s = fma(wa1, J1[c], fma(wa0, J0[c], s));
            }
            acc[22 + a] = fma(wa1, r1, fma(wa0, r0, acc[22 + a]));
        }
    }
    if (LT == TR_LOSS_CAUCHY) acc[0] += log(pm) + (double)pe * 0.69314718055994530942;
}

// One chunk of a pass at the pose in LDS (q = pose[0..3], t = pose[4..6]):
// the robust cost (jaco.cuh:31-104, points with z >= 0) and the normal
// equations (jaco.cuh:106-286) for R <- R exp([w]x), t <- t + R dt over
// inliers [p0, p1).  The 2x6 Jacobian of a point is [D (e_j x X) | D] with
// D = dproj/dZ * R.  The 28 sums land in out[] (LDS) for every thread.
// Multiply-adds may fuse here: the pass differs from the host oracle by
// rounding only.  Called from one site of k_lm_iter, so every chunk of
// every path runs the same instructions.
__device__ __forceinline__ void lm_chunk_sums(const double* __restrict__ lx,
                                              const double* __restrict__ lX,
                                              const double* __restrict__ lw, int p0, int p1,
                                              Cam4 K, LmLoss L, const double* pose, double* out,
                                              double* s_red) {
    const double q[4] = {pose[0], pose[1], pose[2], pose[3]};
    const double t0 = pose[4], t1 = pose[5], t2 = pose[6];
    double R[9];
    dev_quat_to_rotmat(q, R);
    double acc[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) acc[i] = 0.0;
    switch (L.type) {
        case TR_LOSS_TRIVIAL:
            lm_chunk_loop<TR_LOSS_TRIVIAL>(lx, lX, lw, p0, p1, K, L, R, t0, t1, t2, acc);
            break;
        case TR_LOSS_TRUNCATED:
            lm_chunk_loop<TR_LOSS_TRUNCATED>(lx, lX, lw, p0, p1, K, L, R, t0, t1, t2, acc);
            break;
        case TR_LOSS_HUBER:
            lm_chunk_loop<TR_LOSS_HUBER>(lx, lX, lw, p0, p1, K, L, R, t0, t1, t2, acc);
            break;
        case TR_LOSS_TRUNCATED_LE_ZACH:
            lm_chunk_loop<TR_LOSS_TRUNCATED_LE_ZACH>(lx, lX, lw, p0, p1, K, L, R, t0, t1, t2,
                                                     acc);
            break;
        default:
            lm_chunk_loop<TR_LOSS_CAUCHY>(lx, lX, lw, p0, p1, K, L, R, t0, t1, t2, acc);
            break;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double v = wave_transpose_reduce(acc);
    const int idx = lane >> 1;
    if (!(lane & 1) && idx < kLmSums) s_red[wave * kLmSums + idx] = v;
    __syncthreads();
    if (threadIdx.x < kLmSums) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kLmThreads / 64; ++w) s += s_red[w * kLmSums + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// LL^T solve reading the lower triangle of A (lm_utils.cuh:14-100):
// non-positive pivots are floored to 1e-15.  Each pivot is inverted once and
// the substitutions multiply by the inverse (6 FP64 divisions instead of
// 27, on the step's single thread): the result differs from the
// reference's divisions by rounding only.
__device__ __forceinline__ void chol6_solve(const double* A, const double* b, double* x) {
    double L[6][6], inv[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < j; ++k) s += L[i][k] * L[j][k];
            if (i == j) {
                double v = A[6 * i + i] - s;
                if (v <= 0.0) v = 1e-15;
                L[i][i] = sqrt(v);
                inv[i] = 1.0 / L[i][i];
            } else {
                L[i][j] = (A[6 * i + j] - s) * inv[j];
            }
        }
    }
    double y[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < i; ++j) s += L[i][j] * y[j];
        y[i] = (b[i] - s) * inv[i];
    }
#pragma unroll
    for (int i = 5; i >= 0; --i) {
        double s = 0.0;
#pragma unroll
        for (int j = i + 1; j < 6; ++j) s += L[j][i] * x[j];
        x[i] = (y[i] - s) * inv[i];
    }
}

__device__ __forceinline__ void unpack_normal(const double* sums, double* JtJ, double* Jtr) {
    for (int a = 0; a < 6; ++a)
        for (int c = 0; c <= a; ++c) JtJ[6 * a + c] = sums[1 + a * (a + 1) / 2 + c];
    for (int a = 0; a < 6; ++a) Jtr[a] = sums[22 + a];
}

// One LM step after pass j (j = 0: the pass at the starting pose) with the
// pass's sums; then, unless the frame stops, the next proposal in S->prop.
// lm.cuh:15-171.  Returns 1 when the frame is done.  The state's core and
// the sums (in LDS) are copied into registers first and the core written
// back at the end: one round of LDS latency instead of one per field
// access.  The statistics go straight to `stats` (LDS).
__device__ __forceinline__ int lm_step_regs(LmCore* S, LmState* stats, const double* sums, int j,
                                            const LmParams& P) {
    if (j == 0) {
        S->cost = S->initial_cost = sums[0];
        unpack_normal(sums, S->JtJ, S->Jtr);
        S->recompute = 1;
    } else {
        if (sums[0] < S->cost) {
            for (int a = 0; a < 7; ++a) S->cur[a] = S->prop[a];
            S->lambda = fmax(P.min_lambda, S->lambda / 10.0);
            S->cost = sums[0];
            S->recompute = 1;
            unpack_normal(sums, S->JtJ, S->Jtr);
        } else {
            for (int a = 0; a < 6; ++a) S->JtJ[7 * a] -= S->lambda;
            S->lambda = fmin(P.max_lambda, S->lambda * 10.0);
            S->recompute = 0;
            ++stats->invalid;
        }
        ++S->iters;
    }
    if (j >= P.max_iterations) return 1;
    int stop = 0;
    if (S->recompute) {
        double g = 0.0;
        for (int a = 0; a < 6; ++a) g += S->Jtr[a] * S->Jtr[a];
        const double gn = sqrt(g);
        stats->grad_norm = gn;
        if (gn < P.gradient_tol) stop = 1;
    }
    for (int a = 0; a < 6; ++a) S->JtJ[7 * a] += S->lambda;
    double sol[6];
    chol6_solve(S->JtJ, S->Jtr, sol);
    double s2 = 0.0;
    for (int a = 0; a < 6; ++a) {
        sol[a] = -sol[a];
        s2 += sol[a] * sol[a];
    }
    const double sn = sqrt(s2);
    stats->step_norm = sn;
    if (sn < P.step_tol) stop = 1;
    const double q[4] = {S->cur[0], S->cur[1], S->cur[2], S->cur[3]};
    double dq[4], qn[4], dt[3];
    dev_quat_exp(sol, dq);
    dev_quat_mul(q, dq, qn);
    dev_quat_rot(q, sol + 3, dt);
    for (int a = 0; a < 4; ++a) S->prop[a] = qn[a];
    for (int a = 0; a < 3; ++a) S->prop[4 + a] = S->cur[4 + a] + dt[a];
    return stop;
}

__device__ int lm_step(LmState* Sp, const double* sums_lds, int j, const LmParams& P) {
    LmCore S = *static_cast<const LmCore*>(Sp);
    double sums[kLmSums];
#pragma unroll
    for (int i = 0; i < kLmSums; ++i) sums[i] = sums_lds[i];
    const int stop = lm_step_regs(&S, Sp, sums, j, P);
    *static_cast<LmCore*>(Sp) = S;
    return stop;
}

__device__ void lm_write_result(FrameState* st, const LmState* S, int n) {
    tr_pose_result& r = st->result;
    for (int a = 0; a < 4; ++a) r.q[a] = S->cur[a];
    for (int a = 0; a < 3; ++a) r.t[a] = S->cur[4 + a];
    r.iterations = st->samples_drawn;
    r.refinements = S->iters;
    r.num_inliers = (uint64_t)S->nin;
    r.inlier_ratio = (double)S->nin / (double)n;
    r.model_score = st->best_id != -1 ? st->best_score : DBL_MAX;
    r.num_candidates = st->total_candidates;
    r.best_hypothesis = st->best_id;
    r.best_inliers = st->best_inliers;
    r.lm_initial_cost = S->initial_cost;
    r.lm_cost = S->cost;
    r.lm_invalid_steps = S->invalid;
    r.lm_lambda = S->lambda;
    r.lm_step_norm = S->step_norm;
    r.lm_grad_norm = S->grad_norm;
    r.status = st->status;
    r.reserved = 0;
    st->n_inliers = (uint32_t)S->nin;
}

// Compaction, one workgroup per block of kPtBlock points: the block's
// offset is the sum of the counts of the frame's earlier blocks, and the
// inliers keep point order.  The frame's first block also starts the
// frame's LM state at the RANSAC pose.
__global__ __launch_bounds__(kPtBlock) void k_lm_compact(const BatchDev* __restrict__ bp) {
    const BatchDev b = *bp;
    __shared__ uint32_t s_w[kPtBlock / 64];
    __shared__ uint32_t s_w2[kPtBlock / 64];
    const int nb = b.hdr->n_pt_blocks;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int it = blockIdx.x; it < nb; it += gridDim.x) {
        const BlockItem bi = b.pt_blocks[it];
        const FrameDesc& F = b.frames[bi.frame];
        if (bi.start >= F.n) continue;  // padding blocks (uniform)
        const int blk = bi.start / kPtBlock;
        // Sum of the earlier blocks' counts and, for block 0, of all blocks.
        uint32_t before = 0, total = 0;
        for (int k = tid; k < F.mask_blocks; k += kPtBlock) {
            const uint32_t c = F.blk_count[k];
            before += (k < blk) ? c : 0u;
            total += c;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            before += __shfl_xor(before, off, 64);
            total += __shfl_xor(total, off, 64);
        }
        const int i = bi.start + tid;
        const bool in = i < F.n && F.mask[i];
        const uint64_t bal = __ballot(in);
        if (lane == 0) {
            s_w[wave] = before;
            s_w2[wave] = total;
        }
        __syncthreads();
        uint32_t base = 0, nin = 0;
#pragma unroll
        for (int w = 0; w < kPtBlock / 64; ++w) {
            base += s_w[w];
            nin += s_w2[w];
        }
        __syncthreads();
        if (lane == 0) s_w[wave] = (uint32_t)__popcll(bal);
        __syncthreads();
        for (int w = 0; w < wave; ++w) base += s_w[w];
        if (in) {
            const uint32_t o = base + (uint32_t)__popcll(bal & lt);
            const double sc = F.inv_f;
            F.lm_x[2 * o] = F.uv[2 * i] * sc;
            F.lm_x[2 * o + 1] = F.uv[2 * i + 1] * sc;
            F.lm_X[3 * o] = F.xyz[3 * i];
            F.lm_X[3 * o + 1] = F.xyz[3 * i + 1];
            F.lm_X[3 * o + 2] = F.xyz[3 * i + 2];
            if (F.lm_w) F.lm_w[o] = F.w[i];
        }
        if (blk == 0 && tid == 0) {
            const FrameState* st = F.state;
            LmState* S = F.lm_state;
            for (int a = 0; a < 4; ++a) S->cur[a] = S->prop[a] = st->mask_q[a];
            for (int a = 0; a < 3; ++a) S->cur[4 + a] = S->prop[4 + a] = st->mask_t[a];
            S->cost = S->initial_cost = 0.0;
            S->lambda = F.lm.initial_lambda;
            S->grad_norm = S->step_norm = -1.0;
            S->invalid = 0;
            S->iters = 0;
            S->nin = (int32_t)nin;
            const int chunks = nin ? (int)((nin + kLmChunk - 1) / kLmChunk) : 1;
            S->chunks = chunks;
            S->arrived = 0;
            S->recompute = 1;
            S->done = 0;
            // The frame's LM items: one per chunk of its inliers.  Their
            // order in the list does not matter: every form of the pass
            // folds a frame's chunk sums in chunk order.
            const uint32_t at = atomicAdd(&b.hdr->n_lm_live, (uint32_t)chunks);
            for (int c = 0; c < chunks; ++c) b.lm_items[at + c] = {bi.frame, c};
        }
        __syncthreads();
    }
}

// LM launch j of a batch.  Items are (frame, chunk); chunks past the
// frame's inliers, and frames already done, are skipped.  In the final
// launch a frame that needs more passes than this one is finished by its
// chunk-0 item alone.  The step runs on a copy of the frame's state in LDS.
__global__ __launch_bounds__(kLmThreads) void k_lm_iter(const BatchDev* __restrict__ bp, int j,
                                                        int final_launch) {
    static_assert(sizeof(LmState) % 8 == 0 && sizeof(LmState) / 8 <= kLmThreads, "LmState copy");
    constexpr int kWords = sizeof(LmState) / 8;
    const BatchDev b = *bp;
    __shared__ double s_red[(kLmThreads / 64) * kLmSums];
    __shared__ double s_cs[kLmSums];   // sums of the chunk just evaluated
    __shared__ double s_tot[kLmSums];  // sums of the whole pass
    __shared__ double s_pose[7];
    __shared__ LmState s_S;
    __shared__ int s_flag;
    const int tid = threadIdx.x;
    const int ni = (int)b.hdr->n_lm_live;
    for (int it = blockIdx.x; it < ni; it += gridDim.x) {
        const BlockItem bi = b.lm_items[it];
        const FrameDesc& F = b.frames[bi.frame];
        LmState* S = F.lm_state;
        const bool solo = final_launch && j < F.lm.max_iterations;
        const LmParams P = F.lm;  // in registers before the step needs them
        if (tid == 0) {
            int go = !S->done && bi.start < S->chunks;
            if (solo) go = go && bi.start == 0;
            if (go) {
                double p[7];
                for (int a = 0; a < 7; ++a) p[a] = S->prop[a];
                for (int a = 0; a < 7; ++a) s_pose[a] = p[a];
            }
            s_flag = go;
        }
        __syncthreads();
        const int go = s_flag;
        __syncthreads();
        if (!go) continue;

        const int nin = S->nin, chunks = S->chunks;
        const double sc = F.inv_f;
        const Cam4 K = {F.fx * sc, F.fy * sc, F.cx * sc, F.cy * sc};
        LmLoss L;
        L.type = F.lm.loss_type;
        L.s = F.lm.loss_scale;
        L.s2 = L.s * L.s;
        L.inv_s2 = 1.0 / L.s2;
        int c = solo ? 0 : bi.start;
        int jj = j;
        bool staged = false;  // s_S holds the frame's state
        // The frame's state, loaded now so that its latency overlaps the
        // chunk's evaluation; only the frame's last chunk to arrive writes
        // the state, after every chunk has arrived, so this copy is current
        // when that workgroup stages it.
        const uint64_t state_word = tid < kWords ? reinterpret_cast<const uint64_t*>(S)[tid] : 0;
        for (;;) {
            lm_chunk_sums(F.lm_x, F.lm_X, F.lm_w, c * kLmChunk, min(nin, (c + 1) * kLmChunk), K,
                          L, s_pose, s_cs, s_red);
            if (!solo) {
                // Publish the chunk; the last chunk of the frame to arrive
                // folds.  The sums travel through read-modify-write atomics,
                // which take effect at the device's point of coherence and
                // return only after they did, so no cache-wide fence is
                // needed: a chunk's exchanges complete before its arrival
                // is counted, and the folding reads come after the count.
                if (tid < kLmSums) {
                    const uint64_t old = atomicExch(
                        reinterpret_cast<unsigned long long*>(F.lm_part + (size_t)c * kLmSums + tid),
                        (unsigned long long)__double_as_longlong(s_cs[tid]));
                    s_cs[tid] = __longlong_as_double((long long)old);  // waits for the exchange
                }
                __syncthreads();
                if (tid == 0) s_flag = atomicAdd(&S->arrived, 1u) == (uint32_t)chunks - 1u;
                __syncthreads();
                const int last = s_flag;
                __syncthreads();
                if (!last) break;
                if (tid < kLmSums) {
                    unsigned long long* part = reinterpret_cast<unsigned long long*>(F.lm_part);
                    double v[kLmMaxFold];
                    double s = 0.0;
                    for (int c0 = 0; c0 < chunks; c0 += kLmMaxFold) {
                        const int m = min(kLmMaxFold, chunks - c0);
#pragma unroll
                        for (int k = 0; k < kLmMaxFold; ++k)
                            if (k < m)
                                v[k] = __longlong_as_double(
                                    (long long)atomicOr(part + (size_t)(c0 + k) * kLmSums + tid, 0ull));
#pragma unroll
                        for (int k = 0; k < kLmMaxFold; ++k)
                            if (k < m) s = (c0 + k) ? s + v[k] : v[k];
                    }
                    s_tot[tid] = s;
                }
            } else {
                if (tid < kLmSums) s_tot[tid] = c ? s_tot[tid] + s_cs[tid] : s_cs[tid];
                if (c + 1 < chunks) {
                    ++c;
                    continue;
                }
            }
            if (!staged && tid < kWords) reinterpret_cast<uint64_t*>(&s_S)[tid] = state_word;
            staged = true;
            __syncthreads();
            if (tid == 0) {
                s_S.arrived = 0;
                const int done = lm_step(&s_S, s_tot, jj, P);
                s_S.done = done;
                if (done) lm_write_result(F.state, &s_S, F.n);
                for (int a = 0; a < 7; ++a) s_pose[a] = s_S.prop[a];
                s_flag = done;
            }
            __syncthreads();
            const int done = s_flag;
            if (!solo || done) {
                if (tid < kWords)
                    reinterpret_cast<uint64_t*>(S)[tid] = reinterpret_cast<const uint64_t*>(&s_S)[tid];
                __syncthreads();
                break;
            }
            __syncthreads();
            ++jj;
            c = 0;
        }
    }
}

// Refinement on given correspondences (bundle_adjust_cuda, ba.cu:30-130):
// one thread per frame starts the frame's LM state at F.init with all F.n
// correspondences, which the host placed in lm_x / lm_X (lm_w), resets the
// fields of the frame's state that the result reads, and lists the frame's
// LM items.  Frames follow one another in the item list.
__global__ void k_lm_start(const BatchDev* __restrict__ bp) {
    const BatchDev b = *bp;
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t at = 0;
    for (int f = 0; f < b.hdr->n_frames; ++f) {
        const FrameDesc& F = b.frames[f];
        FrameState* st = F.state;
        st->samples_drawn = 0;
        st->total_candidates = 0;
        st->best_id = -1;
        st->best_score = DBL_MAX;
        st->best_inliers = 0;
        st->status = 0;
        LmState* S = F.lm_state;
        for (int a = 0; a < 7; ++a) S->cur[a] = S->prop[a] = F.init[a];
        S->cost = S->initial_cost = 0.0;
        S->lambda = F.lm.initial_lambda;
        S->grad_norm = S->step_norm = -1.0;
        S->invalid = 0;
        S->iters = 0;
        S->nin = F.n;
        const int chunks = F.n ? (F.n + kLmChunk - 1) / kLmChunk : 1;
        S->chunks = chunks;
        S->arrived = 0;
        S->recompute = 1;
        S->done = 0;
        for (int c = 0; c < chunks; ++c) b.lm_items[at + c] = {f, c};
        at += (uint32_t)chunks;
    }
    b.hdr->n_lm_live = at;
}

// get_inliers_mx (entry.cu:12-26): FP64, the divisions of hnormalized.
__global__ __launch_bounds__(kPtBlock) void k_inliers(PoseArg pose, const double* __restrict__ x,
                                                      const double* __restrict__ X, int64_t n,
                                                      double thr, uint8_t* __restrict__ out,
                                                      unsigned long long* __restrict__ count) {
    double R[9];
    dev_quat_to_rotmat(pose.q, R);
    for (int64_t i = blockIdx.x * (int64_t)kPtBlock + threadIdx.x; i - threadIdx.x < n;
         i += (int64_t)gridDim.x * kPtBlock) {
        bool in = false;
        if (i < n) {
            const double X0 = X[3 * i], X1 = X[3 * i + 1], X2 = X[3 * i + 2];
            const double Z0 = R[0] * X0 + R[1] * X1 + R[2] * X2 + pose.t[0];
            const double Z1 = R[3] * X0 + R[4] * X1 + R[5] * X2 + pose.t[1];
            const double Z2 = R[6] * X0 + R[7] * X1 + R[8] * X2 + pose.t[2];
            const double r0 = Z0 / Z2 - x[2 * i];
            const double r1 = Z1 / Z2 - x[2 * i + 1];
            in = r0 * r0 + r1 * r1 < thr && Z2 > 0.0;
            out[i] = in ? 1 : 0;
        }
        const uint64_t bal = __ballot(in);
        if ((threadIdx.x & 63) == 0 && bal) atomicAdd(count, (unsigned long long)__popcll(bal));
    }
}

// ---------------------------------------------------------------------------
// Sampler alone (stage access for tests).
// ---------------------------------------------------------------------------
__global__ void k_sample_only(uint64_t seed, uint32_t n, uint64_t first, int64_t count,
                              int32_t* out) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < count;
         s += (int64_t)gridDim.x * blockDim.x) {
        int32_t idx[3];
        sample_triple(seed, n, first + (uint64_t)s, idx);
        out[3 * s] = idx[0];
        out[3 * s + 1] = idx[1];
        out[3 * s + 2] = idx[2];
    }
}

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
hipError_t launch_prepare(const BatchDev* b, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_prepare, dim3(grid), dim3(kPtBlock), 0, st, b);
    return hipGetLastError();
}
hipError_t launch_sample(const BatchDev* b, int grid, int round, hipStream_t st) {
    hipLaunchKernelGGL(k_sample_p3p, dim3(grid), dim3(kSampleBlock), 0, st, b, round);
    return hipGetLastError();
}
hipError_t launch_select(const BatchDev* b, int grid_sum, int grid, int round, hipStream_t st) {
    hipLaunchKernelGGL(k_sumscores, dim3(grid_sum), dim3(kPtBlock), 0, st, b, round);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_select, dim3(grid), dim3(kSelectThreads), 0, st, b, round);
    return hipGetLastError();
}
hipError_t launch_rerank(const BatchDev* b, int grid, int round, hipStream_t st) {
    hipLaunchKernelGGL(k_rerank, dim3(grid), dim3(kRerankThreads), 0, st, b, round);
    return hipGetLastError();
}
hipError_t launch_choose(const BatchDev* b, int grid, int round, hipStream_t st) {
    hipLaunchKernelGGL(k_choose, dim3(grid), dim3(64), 0, st, b, round);
    return hipGetLastError();
}
hipError_t launch_mask(const BatchDev* b, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_mask, dim3(grid), dim3(kPtBlock), 0, st, b);
    return hipGetLastError();
}
hipError_t launch_lm(const BatchDev* b, int grid, int grid_iter, int launches, hipStream_t st) {
    if (launches < 1 || launches > kLmMaxLaunches) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_lm_compact, dim3(grid), dim3(kPtBlock), 0, st, b);
    hipError_t e = hipGetLastError();
    for (int j = 0; j < launches && e == hipSuccess; ++j) {
        hipLaunchKernelGGL(k_lm_iter, dim3(grid_iter), dim3(kLmThreads), 0, st, b, j,
                           (int)(j == launches - 1));
        e = hipGetLastError();
    }
    return e;
}
hipError_t launch_lm_given(const BatchDev* b, int grid_iter, int launches, hipStream_t st) {
    if (launches < 1 || launches > kLmMaxLaunches) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_lm_start, dim3(1), dim3(64), 0, st, b);
    hipError_t e = hipGetLastError();
    for (int j = 0; j < launches && e == hipSuccess; ++j) {
        hipLaunchKernelGGL(k_lm_iter, dim3(grid_iter), dim3(kLmThreads), 0, st, b, j,
                           (int)(j == launches - 1));
        e = hipGetLastError();
    }
    return e;
}
hipError_t launch_inliers(PoseArg pose, const double* x, const double* X, int64_t n, double thr,
                          uint8_t* inliers, unsigned long long* count, hipStream_t st) {
    const int64_t blocks = (n + kPtBlock - 1) / kPtBlock;
    const int grid = (int)(blocks < 8192 ? (blocks > 0 ? blocks : 1) : 8192);
    hipLaunchKernelGGL(k_inliers, dim3(grid), dim3(kPtBlock), 0, st, pose, x, X, n, thr,
                       inliers, count);
    return hipGetLastError();
}
hipError_t launch_sample_only(uint64_t seed, uint32_t n, uint64_t first, int64_t count,
                              int32_t* out, hipStream_t st) {
    const int64_t blocks = (count + 255) / 256;
    const int grid = (int)(blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096);
    hipLaunchKernelGGL(k_sample_only, dim3(grid), dim3(256), 0, st, seed, n, first, count, out);
    return hipGetLastError();
}

// MSAC kernel variants.
// Compiled MSAC variants: (engine, J or M, T, W) with engine 0 = k_msac
// (J hypotheses per lane, LDS tiles of T points), 1 = k_msac_mfma (M tiles
// of 16 hypotheses per wave), 2 = k_msac_gated (J points per lane, T
// hypotheses per work item), 3 = k_msac_gated2 (the same, two hypotheses
// per pass).
#define TR_SCORE_VARIANTS(X) \
    X(0, 1, 512, 4, k_msac)        \
    X(0, 1, 1024, 4, k_msac)       \
    X(0, 2, 512, 4, k_msac)        \
    X(0, 2, 1024, 4, k_msac)       \
    X(0, 4, 1024, 4, k_msac)       \
    X(0, 2, 1024, 8, k_msac)       \
    X(1, 2, 1024, 4, k_msac_mfma)  \
    X(1, 4, 1024, 4, k_msac_mfma)  \
    X(1, 4, 512, 4, k_msac_mfma)   \
    X(1, 8, 1024, 4, k_msac_mfma)  \
    X(1, 4, 1024, 8, k_msac_mfma)  \
    X(2, 8, 32, 4, k_msac_gated)   \
    X(2, 8, 64, 4, k_msac_gated)   \
    X(2, 16, 32, 4, k_msac_gated)  \
    X(2, 16, 64, 4, k_msac_gated)  \
    X(2, 12, 64, 4, k_msac_gated)  \
    X(3, 8, 64, 4, k_msac_gated2)  \
    X(3, 12, 64, 4, k_msac_gated2) \
    X(3, 16, 64, 4, k_msac_gated2) \
    X(3, 16, 32, 4, k_msac_gated2)

#define TR_SHAPE_IS(E_, J_, T_, W_) \
    (s.engine == E_ && s.cand_per_lane == J_ && s.tile_pts == T_ && s.waves == W_)

int score_variant_list(int32_t* ejtw, int cap) {
    int n = 0;
#define TR_LIST(E_, J_, T_, W_, K_)  \
    if (ejtw && n < cap) {           \
        ejtw[4 * n] = E_;            \
        ejtw[4 * n + 1] = J_;        \
        ejtw[4 * n + 2] = T_;        \
        ejtw[4 * n + 3] = W_;        \
    }                                \
    ++n;
    TR_SCORE_VARIANTS(TR_LIST)
#undef TR_LIST
    return n;
}

bool score_shape_supported(const ScoreShape& s) {
#define TR_MATCH(E_, J_, T_, W_, K_) \
    if (TR_SHAPE_IS(E_, J_, T_, W_)) return true;
    TR_SCORE_VARIANTS(TR_MATCH)
#undef TR_MATCH
    return false;
}

hipError_t launch_score(const BatchDev* b, const ScoreShape& s, int grid, int round,
                        hipStream_t st) {
#define TR_LAUNCH(E_, J_, T_, W_, K_)                                                    \
    if (TR_SHAPE_IS(E_, J_, T_, W_)) {                                                   \
        hipLaunchKernelGGL((K_<J_, T_, W_>), dim3(grid), dim3(64 * W_), 0, st, b, round); \
        return hipGetLastError();                                                 \
    }
    TR_SCORE_VARIANTS(TR_LAUNCH)
#undef TR_LAUNCH
    return hipErrorInvalidValue;
}

int score_blocks_per_cu(const ScoreShape& s) {
    int nb = 0;
#define TR_OCC(E_, J_, T_, W_, K_)                                                        \
    if (TR_SHAPE_IS(E_, J_, T_, W_)) {                                                    \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, K_<J_, T_, W_>, 64 * W_, 0) != \
            hipSuccess)                                                                    \
            nb = 0;                                                                        \
        return nb;                                                                         \
    }
    TR_SCORE_VARIANTS(TR_OCC)
#undef TR_OCC
    return 0;
}
#undef TR_SHAPE_IS

}  // namespace tr
