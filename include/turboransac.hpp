// turboransac.hpp - header-only C++ wrapper over the C ABI (turboransac.h)
// with the shapes of the reference's public calls
// (Module/TbrCudaPnP/include/TbrCudaPnP/entry.h:13-25, ba.cuh:17-24):
//
//   RansacStats estimate_absolute_pose(points2D, points3D, camera,
//                                      ransac_opt, bundle_opt, &pose, &inliers);
//   void get_inliers(pose, x, X, sq_threshold, &inliers);
//   BundleStats bundle_adjust(points_num, x, X, camera, &pose, poses_num,
//                             opt, weights);
//
// Same units and conventions: pixels and world points in, world-to-camera
// pose q = [w, x, y, z], t out, inlier mask of the RANSAC pose as
// std::vector<char>.  Errors throw turboransac::Error (the reference exits
// the process, cudaCheck.h:4-14).  With Eigen available, overloads take
// std::vector<Eigen::Vector2d/3d> like the reference.
#pragma once

#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "turboransac.h"

#if defined(__has_include)
#if __has_include(<Eigen/Dense>)
#include <Eigen/Dense>
#define TURBORANSAC_HAVE_EIGEN 1
#endif
#endif

namespace turboransac {

struct Vec2d {
    double x, y;
};
struct Vec3d {
    double x, y, z;
};

class Error : public std::runtime_error {
  public:
    Error(tr_status s, const std::string& msg)
        : std::runtime_error(std::string(tr_status_string(s)) + ": " + msg), status(s) {}
    tr_status status;
};

// Camera_cuda (camera.cuh:10-85): PINHOLE, params = fx, fy, cx, cy.
struct Camera {
    int model_id = 1;
    int width = -1;
    int height = -1;
    double params[4] = {1.0, 1.0, 0.0, 0.0};
    double focal() const { return 0.5 * (params[0] + params[1]); }
    tr_camera c() const {
        tr_camera o{};
        o.model_id = model_id;
        o.width = width;
        o.height = height;
        for (int i = 0; i < 4; ++i) o.params[i] = params[i];
        return o;
    }
};

// RansacOptions_cuda / BundleOptions_cuda with the reference's defaults.
struct RansacOptions : tr_ransac_options {
    RansacOptions() { tr_ransac_options_default(this); }
};
struct BundleOptions : tr_bundle_options {
    BundleOptions() { tr_bundle_options_default(this); }
};

// RansacStats_cuda (utils.h:25-32), filled in.
struct RansacStats {
    size_t refinements = 0;
    size_t iterations = 0;
    size_t num_inliers = 0;
    double inlier_ratio = 0.0;
    double model_score = 0.0;
};

// BundleStats_cuda (utils.h:62-72), filled in.
struct BundleStats {
    size_t iterations = 0;
    double initial_cost = 0.0;
    double cost = 0.0;
    double lambda = 0.0;
    size_t invalid_steps = 0;
    double step_norm = -1.0;
    double grad_norm = -1.0;
};

// Statistics of a fundamental-matrix estimate (no counterpart in the
// reference checkout).
struct FundamentalStats {
    uint64_t iterations = 0;     // 7-point samples drawn
    uint64_t refinements = 0;    // IRLS iterations
    uint64_t num_inliers = 0;
    double inlier_ratio = 0.0;
    double model_score = 0.0;    // MSAC score of the RANSAC model, pixels^2
    int64_t best_hypothesis = -1;
};

// CameraPose_cuda (pose.cuh:12-80).
struct CameraPose {
    double q[4] = {1.0, 0.0, 0.0, 0.0};
    double t[3] = {0.0, 0.0, 0.0};
    // Row-major rotation of q.
    void R(double out[9]) const {
        const double w = q[0], x = q[1], y = q[2], z = q[3];
        out[0] = 1 - 2 * (y * y + z * z); out[1] = 2 * (x * y - z * w); out[2] = 2 * (x * z + y * w);
        out[3] = 2 * (x * y + z * w); out[4] = 1 - 2 * (x * x + z * z); out[5] = 2 * (y * z - x * w);
        out[6] = 2 * (x * z - y * w); out[7] = 2 * (y * z + x * w); out[8] = 1 - 2 * (x * x + y * y);
    }
};

// Owns a tr_context (one stream and reusable device buffers).
class Engine {
  public:
    explicit Engine(int device = 0) {
        const tr_status s = tr_context_create(device, &ctx_);
        if (s != TR_OK) throw Error(s, "tr_context_create");
    }
    ~Engine() { tr_context_destroy(ctx_); }
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    tr_context* handle() const { return ctx_; }

    RansacStats estimate_absolute_pose(const std::vector<Vec2d>& points2D,
                                       const std::vector<Vec3d>& points3D, const Camera& camera,
                                       const RansacOptions& ransac_opt,
                                       const BundleOptions& bundle_opt, CameraPose* pose,
                                       std::vector<char>* inliers) {
        static_assert(sizeof(Vec2d) == 2 * sizeof(double), "Vec2d layout");
        static_assert(sizeof(Vec3d) == 3 * sizeof(double), "Vec3d layout");
        if (points2D.size() != points3D.size())
            throw Error(TR_ERR_INVALID_ARGUMENT, "2D and 3D point counts must match");
        return run(reinterpret_cast<const double*>(points2D.data()),
                   reinterpret_cast<const double*>(points3D.data()), (int64_t)points2D.size(),
                   camera, ransac_opt, bundle_opt, pose, inliers);
    }

#ifdef TURBORANSAC_HAVE_EIGEN
    RansacStats estimate_absolute_pose(const std::vector<Eigen::Vector2d>& points2D,
                                       const std::vector<Eigen::Vector3d>& points3D,
                                       const Camera& camera, const RansacOptions& ransac_opt,
                                       const BundleOptions& bundle_opt, CameraPose* pose,
                                       std::vector<char>* inliers) {
        if (points2D.size() != points3D.size())
            throw Error(TR_ERR_INVALID_ARGUMENT, "2D and 3D point counts must match");
        std::vector<double> a(2 * points2D.size()), b(3 * points3D.size());
        for (size_t i = 0; i < points2D.size(); ++i) {
            a[2 * i] = points2D[i].x();
            a[2 * i + 1] = points2D[i].y();
            b[3 * i] = points3D[i].x();
            b[3 * i + 1] = points3D[i].y();
            b[3 * i + 2] = points3D[i].z();
        }
        return run(a.data(), b.data(), (int64_t)points2D.size(), camera, ransac_opt, bundle_opt,
                   pose, inliers);
    }
#endif

    // bundle_adjust_cuda (ba.cuh:17-24): refines pose[0] on x (in the
    // camera's pixel units) and X; poses_num is accepted for the
    // reference's signature and must be at least 1.
    BundleStats bundle_adjust(int points_num, const Vec2d* x, const Vec3d* X, const Camera& camera,
                              CameraPose* pose, int poses_num, const BundleOptions& opt,
                              const double* weights = nullptr) {
        if (!pose || poses_num < 1) throw Error(TR_ERR_INVALID_ARGUMENT, "no pose to refine");
        tr_camera cam = camera.c();
        tr_bundle_stats st{};
        const tr_status s = tr_bundle_adjust(ctx_, reinterpret_cast<const double*>(x),
                                             reinterpret_cast<const double*>(X), points_num, &cam,
                                             &opt, weights, pose->q, pose->t, &st);
        if (s != TR_OK) throw Error(s, tr_context_last_error(ctx_));
        BundleStats b;
        b.iterations = st.iterations;
        b.initial_cost = st.initial_cost;
        b.cost = st.cost;
        b.lambda = st.lambda;
        b.invalid_steps = st.invalid_steps;
        b.step_norm = st.step_norm;
        b.grad_norm = st.grad_norm;
        return b;
    }

    // get_inliers_mx (entry.h:22-25): x in normalised image coordinates.
    void get_inliers(const CameraPose& pose, const std::vector<Vec2d>& x,
                     const std::vector<Vec3d>& X, double sq_threshold, std::vector<char>* inliers) {
        if (x.size() != X.size()) throw Error(TR_ERR_INVALID_ARGUMENT, "x and X sizes differ");
        std::vector<uint8_t> m(x.size());
        const tr_status s = tr_get_inliers(ctx_, pose.q, pose.t, reinterpret_cast<const double*>(x.data()),
                                           reinterpret_cast<const double*>(X.data()),
                                           (int64_t)x.size(), sq_threshold, m.data(), nullptr);
        if (s != TR_OK) throw Error(s, tr_context_last_error(ctx_));
        if (inliers) inliers->assign(m.begin(), m.end());
    }

    // The fundamental matrix of pixel correspondences x1[i] <-> x2[i]
    // (tr_estimate_fundamental; no counterpart in the reference checkout):
    // F (row-major, x2^T F x1 = 0) into F[9], the stats returned, the mask
    // of the RANSAC model into *inliers.
    FundamentalStats estimate_fundamental(const std::vector<Vec2d>& x1,
                                          const std::vector<Vec2d>& x2,
                                          const RansacOptions& ransac_opt,
                                          const BundleOptions& bundle_opt, double F[9],
                                          std::vector<char>* inliers) {
        if (x1.size() != x2.size()) throw Error(TR_ERR_INVALID_ARGUMENT, "x1 and x2 sizes differ");
        std::vector<uint8_t> m(inliers ? x1.size() : 0);
        tr_fundamental_result r{};
        const tr_status s = tr_estimate_fundamental(
            ctx_, reinterpret_cast<const double*>(x1.data()),
            reinterpret_cast<const double*>(x2.data()), (int64_t)x1.size(), &ransac_opt,
            &bundle_opt, &r, inliers ? m.data() : nullptr);
        if (s != TR_OK) throw Error(s, tr_context_last_error(ctx_));
        if (F)
            for (int i = 0; i < 9; ++i) F[i] = r.F[i];
        if (inliers) inliers->assign(m.begin(), m.end());
        FundamentalStats st;
        st.iterations = r.iterations;
        st.refinements = r.refinements;
        st.num_inliers = r.num_inliers;
        st.inlier_ratio = r.inlier_ratio;
        st.model_score = r.model_score;
        st.best_hypothesis = r.best_hypothesis;
        return st;
    }

    // One two-view frame of a batched call: its correspondences (held by
    // the caller for the call) and options.
    struct PairFrame {
        const std::vector<Vec2d>* x1 = nullptr;
        const std::vector<Vec2d>* x2 = nullptr;
        RansacOptions ransac_opt;
        BundleOptions bundle_opt;
    };

    // Many two-view frames in one call (tr_estimate_fundamental_batch; no
    // counterpart in the reference checkout): frame f's F into (*F)[f], its
    // stats at [f] of the result, its mask into (*inliers)[f].  Each result
    // is bit-identical to the frame's estimate_fundamental call.
    std::vector<FundamentalStats> estimate_fundamental_batch(
        const std::vector<PairFrame>& frames, std::vector<std::array<double, 9>>* F,
        std::vector<std::vector<char>>* inliers) {
        const size_t nf = frames.size();
        std::vector<tr_pair_frame> in(nf);
        std::vector<std::vector<uint8_t>> m(inliers ? nf : 0);
        std::vector<uint8_t*> mp(inliers ? nf : 0);
        for (size_t f = 0; f < nf; ++f) {
            const PairFrame& p = frames[f];
            if (!p.x1 || !p.x2 || p.x1->size() != p.x2->size())
                throw Error(TR_ERR_INVALID_ARGUMENT,
                            "frame " + std::to_string(f) + ": x1 and x2 missing or of different sizes");
            in[f].points1 = reinterpret_cast<const double*>(p.x1->data());
            in[f].points2 = reinterpret_cast<const double*>(p.x2->data());
            in[f].num_points = (int64_t)p.x1->size();
            in[f].ransac = p.ransac_opt;
            in[f].bundle = p.bundle_opt;
            if (inliers) {
                m[f].resize(p.x1->size());
                mp[f] = m[f].data();
            }
        }
        std::vector<tr_fundamental_result> r(nf);
        const tr_status s = tr_estimate_fundamental_batch(ctx_, in.data(), (int32_t)nf, r.data(),
                                                          inliers ? mp.data() : nullptr);
        if (s != TR_OK) throw Error(s, tr_context_last_error(ctx_));
        if (F) F->resize(nf);
        if (inliers) inliers->resize(nf);
        std::vector<FundamentalStats> out(nf);
        for (size_t f = 0; f < nf; ++f) {
            if (F)
                for (int i = 0; i < 9; ++i) (*F)[f][i] = r[f].F[i];
            if (inliers) (*inliers)[f].assign(m[f].begin(), m[f].end());
            out[f].iterations = r[f].iterations;
            out[f].refinements = r[f].refinements;
            out[f].num_inliers = r[f].num_inliers;
            out[f].inlier_ratio = r[f].inlier_ratio;
            out[f].model_score = r[f].model_score;
            out[f].best_hypothesis = r[f].best_hypothesis;
        }
        return out;
    }

  private:
    RansacStats run(const double* p2, const double* p3, int64_t n, const Camera& camera,
                    const RansacOptions& ro, const BundleOptions& bo, CameraPose* pose,
                    std::vector<char>* inliers) {
        tr_camera cam = camera.c();
        tr_pose_result r{};
        // score_initial_model: *pose is the initial model (in/out, as the
        // reference's pose argument).
        if (pose && ro.score_initial_model) {
            for (int i = 0; i < 4; ++i) r.q[i] = pose->q[i];
            for (int i = 0; i < 3; ++i) r.t[i] = pose->t[i];
        }
        std::vector<uint8_t> mask(inliers ? (size_t)n : 0);
        const tr_status s = tr_estimate_absolute_pose(ctx_, p2, p3, n, &cam, &ro, &bo, &r,
                                                      inliers ? mask.data() : nullptr);
        if (s != TR_OK) throw Error(s, tr_context_last_error(ctx_));
        if (pose) {
            for (int i = 0; i < 4; ++i) pose->q[i] = r.q[i];
            for (int i = 0; i < 3; ++i) pose->t[i] = r.t[i];
        }
        if (inliers) inliers->assign(mask.begin(), mask.end());
        RansacStats st;
        st.refinements = r.refinements;
        st.iterations = r.iterations;
        st.num_inliers = r.num_inliers;
        st.inlier_ratio = r.inlier_ratio;
        st.model_score = r.model_score;
        return st;
    }
    tr_context* ctx_ = nullptr;
};

// Free functions with the reference's signatures; they use one engine per
// host thread on device 0.
inline Engine& thread_engine() {
    thread_local Engine engine(0);
    return engine;
}

inline RansacStats estimate_absolute_pose(const std::vector<Vec2d>& points2D,
                                          const std::vector<Vec3d>& points3D,
                                          const Camera& camera, const RansacOptions& ransac_opt,
                                          const BundleOptions& bundle_opt, CameraPose* pose,
                                          std::vector<char>* inliers) {
    return thread_engine().estimate_absolute_pose(points2D, points3D, camera, ransac_opt,
                                                  bundle_opt, pose, inliers);
}

inline void get_inliers(const CameraPose& pose, const std::vector<Vec2d>& x,
                        const std::vector<Vec3d>& X, double sq_threshold,
                        std::vector<char>* inliers) {
    thread_engine().get_inliers(pose, x, X, sq_threshold, inliers);
}

inline FundamentalStats estimate_fundamental(const std::vector<Vec2d>& x1,
                                             const std::vector<Vec2d>& x2,
                                             const RansacOptions& ransac_opt,
                                             const BundleOptions& bundle_opt, double F[9],
                                             std::vector<char>* inliers) {
    return thread_engine().estimate_fundamental(x1, x2, ransac_opt, bundle_opt, F, inliers);
}

inline std::vector<FundamentalStats> estimate_fundamental_batch(
    const std::vector<Engine::PairFrame>& frames, std::vector<std::array<double, 9>>* F,
    std::vector<std::vector<char>>* inliers) {
    return thread_engine().estimate_fundamental_batch(frames, F, inliers);
}

inline BundleStats bundle_adjust(int points_num, const Vec2d* x, const Vec3d* X,
                                 const Camera& camera, CameraPose* pose, int poses_num,
                                 const BundleOptions& opt, const double* weights = nullptr) {
    return thread_engine().bundle_adjust(points_num, x, X, camera, pose, poses_num, opt, weights);
}

}  // namespace turboransac
