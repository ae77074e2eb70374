// example_main.cpp - a C++ caller of the wrapper (include/turboransac.hpp),
// written the way the reference's evaluator calls estimate_absolute_pose_cuda
// (TbrMain/src/main.cu:160-202): build a camera, set the options, call, and
// measure the pose error against the ground truth.
//
//   tr_example --abi            print the ABI version (no GPU needed)
//   tr_example --frame IN OUT   one frame from the binary file IN through
//                               turboransac::estimate_absolute_pose, the
//                               pose, stats and mask written to OUT
//   tr_example --fundamental IN OUT
//                               one two-view frame from IN through
//                               turboransac::estimate_fundamental, F, the
//                               stats and the mask written to OUT
//   tr_example --fundamental-batch IN OUT
//                               many two-view frames (a frame count, then
//                               each frame as for --fundamental) through
//                               turboransac::estimate_fundamental_batch,
//                               each frame's record as for --fundamental
//   tr_example [N] [inliers]    synthetic frame, estimate, print errors
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "turboransac.hpp"

// IN: int64 N, double fx fy cx cy, threshold (px), uint64 samples, seed,
// LM iterations, then N x 2 pixels and N x 3 world points (doubles).
// OUT: q[4], t[3] (doubles), refinements, iterations, num_inliers (uint64),
// inlier_ratio, model_score (doubles), then the N mask bytes.
static int run_frame_file(const char* in_path, const char* out_path) {
    std::FILE* f = std::fopen(in_path, "rb");
    if (!f) return 3;
    int64_t n = 0;
    double K[4], thr;
    uint64_t samples, seed, lm_it;
    bool ok = std::fread(&n, 8, 1, f) == 1 && std::fread(K, 8, 4, f) == 4 &&
              std::fread(&thr, 8, 1, f) == 1 && std::fread(&samples, 8, 1, f) == 1 &&
              std::fread(&seed, 8, 1, f) == 1 && std::fread(&lm_it, 8, 1, f) == 1 && n > 0;
    std::vector<turboransac::Vec2d> p2(ok ? n : 0);
    std::vector<turboransac::Vec3d> p3(ok ? n : 0);
    ok = ok && std::fread(p2.data(), sizeof(turboransac::Vec2d), n, f) == (size_t)n &&
         std::fread(p3.data(), sizeof(turboransac::Vec3d), n, f) == (size_t)n;
    std::fclose(f);
    if (!ok) return 3;
    turboransac::Camera cam;
    for (int i = 0; i < 4; ++i) cam.params[i] = K[i];
    turboransac::RansacOptions ro;
    ro.max_iterations = samples;
    ro.min_iterations = 0;
    ro.adaptive_stopping = 0;
    ro.max_reproj_error = thr;
    ro.seed = seed;
    turboransac::BundleOptions bo;
    bo.max_iterations = lm_it;
    turboransac::CameraPose pose;
    std::vector<char> inliers;
    try {
        const auto st = turboransac::estimate_absolute_pose(p2, p3, cam, ro, bo, &pose, &inliers);
        std::FILE* o = std::fopen(out_path, "wb");
        if (!o) return 3;
        const uint64_t u[3] = {st.refinements, st.iterations, st.num_inliers};
        const double d[2] = {st.inlier_ratio, st.model_score};
        std::fwrite(pose.q, 8, 4, o);
        std::fwrite(pose.t, 8, 3, o);
        std::fwrite(u, 8, 3, o);
        std::fwrite(d, 8, 2, o);
        std::fwrite(inliers.data(), 1, inliers.size(), o);
        std::fclose(o);
        return 0;
    } catch (const turboransac::Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}

// IN: int64 N, double threshold (px), uint64 samples, seed, IRLS
// iterations, double step tolerance, then N x 2 pixels of view 1 and N x 2
// of view 2.  OUT: F[9], refinements, num_inliers (uint64), model_score,
// best_hypothesis (int64), then the N mask bytes.
static int run_fundamental_file(const char* in_path, const char* out_path) {
    std::FILE* f = std::fopen(in_path, "rb");
    if (!f) return 3;
    int64_t n = 0;
    double thr, step_tol;
    uint64_t samples, seed, it;
    bool ok = std::fread(&n, 8, 1, f) == 1 && std::fread(&thr, 8, 1, f) == 1 &&
              std::fread(&samples, 8, 1, f) == 1 && std::fread(&seed, 8, 1, f) == 1 &&
              std::fread(&it, 8, 1, f) == 1 && std::fread(&step_tol, 8, 1, f) == 1 && n > 0;
    std::vector<turboransac::Vec2d> x1(ok ? n : 0), x2(ok ? n : 0);
    ok = ok && std::fread(x1.data(), sizeof(turboransac::Vec2d), n, f) == (size_t)n &&
         std::fread(x2.data(), sizeof(turboransac::Vec2d), n, f) == (size_t)n;
    std::fclose(f);
    if (!ok) return 3;
    turboransac::RansacOptions ro;
    ro.max_iterations = samples;
    ro.max_epipolar_error = thr;
    ro.seed = seed;
    turboransac::BundleOptions bo;
    bo.max_iterations = it;
    bo.step_tol = step_tol;
    double F[9];
    std::vector<char> inliers;
    try {
        const auto st = turboransac::estimate_fundamental(x1, x2, ro, bo, F, &inliers);
        std::FILE* o = std::fopen(out_path, "wb");
        if (!o) return 3;
        const uint64_t u[2] = {st.refinements, st.num_inliers};
        std::fwrite(F, 8, 9, o);
        std::fwrite(u, 8, 2, o);
        std::fwrite(&st.model_score, 8, 1, o);
        std::fwrite(&st.best_hypothesis, 8, 1, o);
        std::fwrite(inliers.data(), 1, inliers.size(), o);
        std::fclose(o);
        return 0;
    } catch (const turboransac::Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}

// One two-view frame of a --fundamental-batch file (the --fundamental
// layout); false at a short or malformed read.
struct PairIn {
    std::vector<turboransac::Vec2d> x1, x2;
    turboransac::RansacOptions ro;
    turboransac::BundleOptions bo;
};
static bool read_pair(std::FILE* f, PairIn* p) {
    int64_t n = 0;
    double thr, step_tol;
    uint64_t samples, seed, it;
    if (!(std::fread(&n, 8, 1, f) == 1 && std::fread(&thr, 8, 1, f) == 1 &&
          std::fread(&samples, 8, 1, f) == 1 && std::fread(&seed, 8, 1, f) == 1 &&
          std::fread(&it, 8, 1, f) == 1 && std::fread(&step_tol, 8, 1, f) == 1 && n > 0 &&
          n <= (int64_t(1) << 28)))
        return false;
    p->x1.resize(n);
    p->x2.resize(n);
    if (std::fread(p->x1.data(), sizeof(turboransac::Vec2d), n, f) != (size_t)n ||
        std::fread(p->x2.data(), sizeof(turboransac::Vec2d), n, f) != (size_t)n)
        return false;
    p->ro.max_iterations = samples;
    p->ro.max_epipolar_error = thr;
    p->ro.seed = seed;
    p->bo.max_iterations = it;
    p->bo.step_tol = step_tol;
    return true;
}

static int run_fundamental_batch_file(const char* in_path, const char* out_path) {
    std::FILE* f = std::fopen(in_path, "rb");
    if (!f) return 3;
    int64_t nf = 0;
    bool ok = std::fread(&nf, 8, 1, f) == 1 && nf > 0 && nf <= 65535;
    std::vector<PairIn> pairs(ok ? nf : 0);
    for (int64_t i = 0; ok && i < nf; ++i) ok = read_pair(f, &pairs[i]);
    std::fclose(f);
    if (!ok) return 3;
    std::vector<turboransac::Engine::PairFrame> frames(nf);
    for (int64_t i = 0; i < nf; ++i) {
        frames[i].x1 = &pairs[i].x1;
        frames[i].x2 = &pairs[i].x2;
        frames[i].ransac_opt = pairs[i].ro;
        frames[i].bundle_opt = pairs[i].bo;
    }
    std::vector<std::array<double, 9>> F;
    std::vector<std::vector<char>> inliers;
    try {
        const auto st = turboransac::estimate_fundamental_batch(frames, &F, &inliers);
        std::FILE* o = std::fopen(out_path, "wb");
        if (!o) return 3;
        for (int64_t i = 0; i < nf; ++i) {
            const uint64_t u[2] = {st[i].refinements, st[i].num_inliers};
            std::fwrite(F[i].data(), 8, 9, o);
            std::fwrite(u, 8, 2, o);
            std::fwrite(&st[i].model_score, 8, 1, o);
            std::fwrite(&st[i].best_hypothesis, 8, 1, o);
            std::fwrite(inliers[i].data(), 1, inliers[i].size(), o);
        }
        std::fclose(o);
        return 0;
    } catch (const turboransac::Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}

int main(int argc, char** argv) {
    if (argc > 1 && std::strcmp(argv[1], "--abi") == 0) {
        std::printf("turboransac abi %d\n", tr_abi_version());
        return 0;
    }
    if (argc > 3 && std::strcmp(argv[1], "--frame") == 0) return run_frame_file(argv[2], argv[3]);
    if (argc > 3 && std::strcmp(argv[1], "--fundamental") == 0)
        return run_fundamental_file(argv[2], argv[3]);
    if (argc > 3 && std::strcmp(argv[1], "--fundamental-batch") == 0)
        return run_fundamental_batch_file(argv[2], argv[3]);
    const int n = argc > 1 ? std::atoi(argv[1]) : 2000;
    const double ratio = argc > 2 ? std::atof(argv[2]) : 0.6;

    // Ground truth: a rotation about a tilted axis, camera 2.5 m from the
    // cloud centre.
    const double ang = 0.4, ax[3] = {0.3, 0.8, 0.52};
    const double an = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    turboransac::CameraPose gt;
    gt.q[0] = std::cos(ang / 2);
    for (int i = 0; i < 3; ++i) gt.q[1 + i] = std::sin(ang / 2) * ax[i] / an;
    double R[9];
    gt.R(R);
    gt.t[0] = 0.1;
    gt.t[1] = -0.2;
    gt.t[2] = 2.5;

    turboransac::Camera cam;
    cam.width = 640;
    cam.height = 480;
    cam.params[0] = cam.params[1] = 525.0;
    cam.params[2] = 320.0;
    cam.params[3] = 240.0;

    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::normal_distribution<double> noise(0.0, 1.0);
    std::vector<turboransac::Vec2d> p2;
    std::vector<turboransac::Vec3d> p3;
    const int n_in = (int)(n * ratio);
    while ((int)p2.size() < n) {
        turboransac::Vec3d X{U(rng), U(rng), U(rng)};
        if ((int)p2.size() < n_in) {
            const double Z0 = R[0] * X.x + R[1] * X.y + R[2] * X.z + gt.t[0];
            const double Z1 = R[3] * X.x + R[4] * X.y + R[5] * X.z + gt.t[1];
            const double Z2 = R[6] * X.x + R[7] * X.y + R[8] * X.z + gt.t[2];
            p2.push_back({525.0 * Z0 / Z2 + 320.0 + noise(rng), 525.0 * Z1 / Z2 + 240.0 + noise(rng)});
        } else {
            p2.push_back({320.0 + 320.0 * U(rng), 240.0 + 240.0 * U(rng)});
        }
        p3.push_back(X);
    }

    turboransac::RansacOptions ro;
    ro.max_iterations = 1024;
    ro.max_reproj_error = 8.0;
    turboransac::BundleOptions bo;
    bo.max_iterations = 25;
    turboransac::CameraPose pose;
    std::vector<char> inliers;
    try {
        const auto t0 = std::chrono::steady_clock::now();
        const auto st = turboransac::estimate_absolute_pose(p2, p3, cam, ro, bo, &pose, &inliers);
        const auto t1 = std::chrono::steady_clock::now();
        double Rp[9];
        pose.R(Rp);
        // Rotation error: angle of R_gt R_pred^T; centre error in cm.
        double tr = 0.0;
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) tr += R[3 * i + k] * Rp[3 * i + k];
        const double rot = std::acos(std::max(-1.0, std::min(1.0, (tr - 1.0) / 2.0))) * 180.0 / M_PI;
        double cg[3], cp[3];
        for (int i = 0; i < 3; ++i) {
            cg[i] = -(R[i] * gt.t[0] + R[3 + i] * gt.t[1] + R[6 + i] * gt.t[2]);
            cp[i] = -(Rp[i] * pose.t[0] + Rp[3 + i] * pose.t[1] + Rp[6 + i] * pose.t[2]);
        }
        const double cm = 100.0 * std::sqrt((cg[0] - cp[0]) * (cg[0] - cp[0]) +
                                            (cg[1] - cp[1]) * (cg[1] - cp[1]) +
                                            (cg[2] - cp[2]) * (cg[2] - cp[2]));
        std::printf("inliers %zu of %d, samples %zu, lm %zu, rot %.5f deg, centre %.5f cm, %.3f ms\n",
                    st.num_inliers, n, st.iterations, st.refinements, rot, cm,
                    std::chrono::duration<double, std::milli>(t1 - t0).count());
        return (rot < 0.5 && cm < 2.0) ? 0 : 2;
    } catch (const turboransac::Error& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
