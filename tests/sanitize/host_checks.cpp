// host_checks.cpp - host code under AddressSanitizer and
// UndefinedBehaviorSanitizer, with no GPU: the oracle's estimators and the
// engine's host-side planning (tuning table parser, boundary validation,
// PROSAC schedule, batch planner and metadata packing).  Built and run by
// tests/test_sanitizers.py; any sanitizer report aborts with a non-zero
// status.
//
//   host_checks oracle    the oracle's estimators on synthetic frames
//   host_checks engine    the engine's host-side planning
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "turboransac.h"

#ifdef TR_CHECK_ORACLE
#include "oracle.h"
#endif
#ifdef TR_CHECK_ENGINE
#include "engine.hpp"
#endif

namespace {

// SplitMix64: a deterministic generator for the synthetic inputs.
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u = uni() + 1e-300, v = uni();
        return std::sqrt(-2.0 * std::log(u)) * std::cos(2.0 * M_PI * v);
    }
};

int failures = 0;

void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        ++failures;
    }
}

// Pixels of a 525 px camera at the identity pose, points 2-6 m ahead;
// a share of the correspondences replaced by uniform outliers.
struct PoseFrame {
    std::vector<double> uv, xyz;
    int64_t n;
};

PoseFrame make_pose_frame(Rng& r, int64_t n, double inlier_ratio) {
    PoseFrame f;
    f.n = n;
    f.uv.resize(2 * n);
    f.xyz.resize(3 * n);
    for (int64_t i = 0; i < n; ++i) {
        const double X = r.uni() * 4.0 - 2.0, Y = r.uni() * 3.0 - 1.5, Z = 2.0 + 4.0 * r.uni();
        f.xyz[3 * i] = X;
        f.xyz[3 * i + 1] = Y;
        f.xyz[3 * i + 2] = Z;
        if (r.uni() < inlier_ratio) {
            f.uv[2 * i] = 525.0 * X / Z + 320.0 + r.normal();
            f.uv[2 * i + 1] = 525.0 * Y / Z + 240.0 + r.normal();
        } else {
            f.uv[2 * i] = 640.0 * r.uni();
            f.uv[2 * i + 1] = 480.0 * r.uni();
        }
    }
    return f;
}

tr_camera camera() {
    tr_camera c{};
    c.model_id = 1;
    c.width = 640;
    c.height = 480;
    c.params[0] = 525.0;
    c.params[1] = 525.0;
    c.params[2] = 320.0;
    c.params[3] = 240.0;
    return c;
}

tr_ransac_options ransac(uint64_t samples, uint64_t seed) {
    tr_ransac_options o{};
    o.max_iterations = samples;
    o.min_iterations = 0;
    o.dyn_num_trials_mult = 3.0;
    o.success_prob = 0.9999;
    o.max_reproj_error = 8.0;
    o.seed = seed;
    o.max_total_iterations = 100000;
    o.adaptive_stopping = 1;
    o.max_prosac_iterations = 100000;
    o.max_epipolar_error = 1.5;
    return o;
}

tr_bundle_options bundle(int loss) {
    tr_bundle_options o{};
    o.max_iterations = 25;
    o.loss_type = loss;
    o.loss_scale = 1.0;
    o.gradient_tol = 1e-10;
    o.step_tol = 1e-8;
    o.initial_lambda = 1e-3;
    o.min_lambda = 1e-10;
    o.max_lambda = 1e10;
    return o;
}

#ifdef TR_CHECK_ORACLE
int check_oracle() {
    Rng r(11);
    const tr_camera cam = camera();
    // per-frame path with the trace, every loss, PROSAC and weights
    for (int loss = TR_LOSS_TRIVIAL; loss <= TR_LOSS_TRUNCATED_LE_ZACH; ++loss) {
        PoseFrame f = make_pose_frame(r, 300 + 97 * loss, 0.6);
        tr_ransac_options ro = ransac(256, 5 + loss);
        ro.progressive_sampling = loss & 1;
        tr_bundle_options bo = bundle(loss);
        tr_pose_result res{};
        std::vector<uint8_t> mask(f.n);
        const int64_t S = 300;
        std::vector<int32_t> samples(3 * S), nsol(S);
        std::vector<double> poses(48 * S), scores(4 * S);
        orc_trace trc{};
        trc.max_samples = S;
        trc.samples = samples.data();
        trc.num_solutions = nsol.data();
        trc.poses = poses.data();
        trc.scores = scores.data();
        const int32_t st = orc_estimate_absolute_pose(f.uv.data(), f.xyz.data(), f.n, &cam, &ro,
                                                      &bo, &res, mask.data(), &trc);
        expect(st == TR_OK, "orc_estimate_absolute_pose status");
        expect(res.num_inliers > (uint64_t)(0.4 * f.n), "pose inliers");
    }
    // tiny frames: 3 points and degenerate (all points equal)
    {
        PoseFrame f = make_pose_frame(r, 3, 1.0);
        tr_ransac_options ro = ransac(64, 1);
        tr_bundle_options bo = bundle(TR_LOSS_CAUCHY);
        tr_pose_result res{};
        std::vector<uint8_t> mask(3);
        expect(orc_estimate_absolute_pose(f.uv.data(), f.xyz.data(), 3, &cam, &ro, &bo, &res,
                                          mask.data(), nullptr) == TR_OK,
               "3-point frame");
        for (int i = 1; i < 3; ++i) {
            f.uv[2 * i] = f.uv[0];
            f.uv[2 * i + 1] = f.uv[1];
            for (int k = 0; k < 3; ++k) f.xyz[3 * i + k] = f.xyz[k];
        }
        expect(orc_estimate_absolute_pose(f.uv.data(), f.xyz.data(), 3, &cam, &ro, &bo, &res,
                                          mask.data(), nullptr) == TR_OK,
               "degenerate frame");
    }
    // frame-parallel estimation on 3 threads with masks
    {
        std::vector<PoseFrame> frames;
        for (int i = 0; i < 6; ++i) frames.push_back(make_pose_frame(r, 200 + 150 * i, 0.5));
        std::vector<tr_frame> fr(frames.size());
        std::vector<std::vector<uint8_t>> masks(frames.size());
        std::vector<uint8_t*> mp(frames.size());
        for (size_t i = 0; i < frames.size(); ++i) {
            fr[i] = tr_frame{};
            fr[i].points2D = frames[i].uv.data();
            fr[i].points3D = frames[i].xyz.data();
            fr[i].num_points = frames[i].n;
            fr[i].camera = cam;
            fr[i].ransac = ransac(128, 40 + i);
            fr[i].bundle = bundle(TR_LOSS_CAUCHY);
            fr[i].initial_q[0] = 1.0;
            masks[i].resize(frames[i].n);
            mp[i] = masks[i].data();
        }
        std::vector<tr_pose_result> res(frames.size());
        expect(orc_estimate_many_masks(fr.data(), (int32_t)fr.size(), 3, res.data(), mp.data()) ==
                   TR_OK,
               "orc_estimate_many_masks");
    }
    // samplers
    {
        std::vector<int32_t> out(7 * 512);
        orc_sample_triples(77, 1000, 0, 512, out.data());
        orc_sample_subset(77, 9, 7, 0, 512, out.data());
        orc_sample_prosac(3, 500, 2000, 3000, 0, 512, out.data());
        tr_ransac_options ro = ransac(100, 0);
        for (double w : {0.0, 1e-9, 0.1, 0.5, 1.0}) (void)orc_needed_samples(w, &ro);
    }
    // cubic roots over six decades of coefficients
    for (int i = 0; i < 2000; ++i) {
        double c[4], x[3];
        for (double& v : c) v = r.normal() * std::pow(10.0, 3.0 * (2.0 * r.uni() - 1.0));
        const int k = orc_solve_cubic(c[0], c[1], c[2], c[3], x);
        expect(k >= 0 && k <= 3, "cubic root count");
    }
    // fundamental matrix: two views, 40% outliers; and all points equal
    {
        const int64_t n = 1500;
        std::vector<double> p1(2 * n), p2(2 * n);
        const double ang = 0.2, tx = 0.4, ty = 0.05, tz = 0.1;
        const double c = std::cos(ang), s = std::sin(ang);
        for (int64_t i = 0; i < n; ++i) {
            const double X = r.uni() * 4.0 - 2.0, Y = r.uni() * 3.0 - 1.5, Z = 3.0 + 4.0 * r.uni();
            p1[2 * i] = 525.0 * X / Z + 320.0;
            p1[2 * i + 1] = 525.0 * Y / Z + 240.0;
            const double X2 = c * X + s * Z + tx, Y2 = Y + ty, Z2 = -s * X + c * Z + tz;
            if (r.uni() < 0.6) {
                p2[2 * i] = 525.0 * X2 / Z2 + 320.0 + 0.5 * r.normal();
                p2[2 * i + 1] = 525.0 * Y2 / Z2 + 240.0 + 0.5 * r.normal();
            } else {
                p2[2 * i] = 640.0 * r.uni();
                p2[2 * i + 1] = 480.0 * r.uni();
            }
        }
        tr_ransac_options ro = ransac(512, 9);
        tr_bundle_options bo = bundle(TR_LOSS_TRIVIAL);
        tr_fundamental_result res{};
        std::vector<uint8_t> mask(n);
        const int64_t S = 64;
        std::vector<int32_t> samples(7 * S), nsol(S);
        std::vector<double> models(27 * S), scores(3 * S);
        orc_ftrace trc{};
        trc.max_samples = S;
        trc.samples = samples.data();
        trc.num_solutions = nsol.data();
        trc.models = models.data();
        trc.scores = scores.data();
        expect(orc_estimate_fundamental(p1.data(), p2.data(), n, &ro, &bo, &res, mask.data(),
                                        &trc) == TR_OK,
               "orc_estimate_fundamental");
        expect(res.num_inliers > (uint64_t)(0.45 * n), "fundamental inliers");
        for (int64_t i = 0; i < 20; ++i) {
            p1[2 * i] = p1[0];
            p1[2 * i + 1] = p1[1];
        }
        expect(orc_estimate_fundamental(p1.data(), p1.data(), 20, &ro, &bo, &res, mask.data(),
                                        nullptr) == TR_OK,
               "degenerate fundamental");
    }
    return failures;
}
#endif

#ifdef TR_CHECK_ENGINE
int check_engine() {
    using namespace tr;
    Rng r(23);
    // tuning table: the built-in one round-trips; malformed lines are errors
    {
        Tuning t = Tuning::defaults();
        Tuning u;
        std::string err;
        expect(u.parse(t.to_text(), "defaults", &err), "tuning round trip");
        const char* bad[] = {"single", "single 10 nosuch 1 1 1 1", "batch -1 valu 1 64 4",
                             "single -1 gated 999 64 4 512", "batch x valu 1 64 4 512",
                             "single -1 valu 1 64 4 100"};
        for (const char* b : bad) {
            Tuning v;
            expect(!v.parse(b, "bad", &err), b);
        }
    }
    // PROSAC schedules, small and large
    {
        std::vector<uint64_t> ev;
        prosac_events(3, 100, 1000, ev);
        prosac_events(10, 100000, 100000, ev);
        prosac_events(200000, 200000, 300000, ev);
        expect(!ev.empty(), "prosac events");
    }
    // boundary validation of good and bad frames
    {
        PoseFrame f = make_pose_frame(r, 50, 0.7);
        const tr_camera cam = camera();
        tr_ransac_options ro = ransac(128, 1);
        tr_bundle_options bo = bundle(TR_LOSS_CAUCHY);
        expect(validate_frame(f.uv.data(), f.xyz.data(), f.n, &cam, &ro, &bo, true).ok(),
               "valid frame");
        f.uv[7] = NAN;
        expect(validate_frame(f.uv.data(), f.xyz.data(), f.n, &cam, &ro, &bo, true).code ==
                   TR_ERR_NOT_FINITE,
               "NaN frame");
        expect(validate_frame(f.uv.data(), f.xyz.data(), 2, &cam, &ro, &bo, true).code ==
                   TR_ERR_TOO_FEW_POINTS,
               "2 points");
        // The one-pass copy and finite check of host inputs.
        {
            const double vals[] = {0.0, -0.0, 1.0, -3.5e300, DBL_MAX, -DBL_MAX, DBL_MIN,
                                   4.9e-324, 123.25};
            std::vector<double> src(vals, vals + 9), dst(9, 7.0);
            expect(copy_finite(dst.data(), src.data(), src.size()) && dst == src, "finite copy");
            for (double bad : {(double)NAN, (double)INFINITY, -(double)INFINITY}) {
                for (size_t at : {size_t(0), size_t(4), size_t(8)}) {
                    std::vector<double> b = src;
                    b[at] = bad;
                    std::vector<double> d(9, 0.0);
                    expect(!copy_finite(d.data(), b.data(), b.size()), "non-finite value");
                    expect(std::memcmp(d.data(), b.data(), sizeof(double) * 9) == 0, "copied");
                }
            }
            expect(copy_finite(dst.data(), src.data(), 0), "empty copy");
        }
        // The largest frame the kernels' 32-bit index products allow:
        // below 2^28 correspondences (device arrays: the values are not read).
        expect(validate_frame(f.uv.data(), f.xyz.data(), (int64_t(1) << 28) - 1, &cam, &ro, &bo,
                              false)
                   .ok(),
               "2^28 - 1 points");
        expect(validate_frame(f.uv.data(), f.xyz.data(), int64_t(1) << 28, &cam, &ro, &bo, false)
                       .code == TR_ERR_CAPACITY,
               "2^28 points");
        tr_bundle_options bad = bo;
        bad.loss_type = 9;
        expect(!validate_bundle_options(&bad).ok(), "bad loss");
    }
    // batch plans: many tiny frames, mixed sizes, PROSAC frames, one-frame
    // plans within reserved capacity; metadata packed into host buffers
    {
        const Tuning tune = Tuning::defaults();
        const int configs[][2] = {{300, 3}, {300, 10}, {64, 5000}, {256, 0}, {1, 200000}};
        for (const auto& cfg : configs) {
            const int nf = cfg[0];
            std::vector<FrameInput> in(nf);
            for (int f = 0; f < nf; ++f) {
                FrameInput& fi = in[f];
                fi.n = cfg[1] ? cfg[1] : 500 + (int64_t)(r.uni() * 20000.0);
                fi.uv = reinterpret_cast<const double*>(uintptr_t(0x10000));
                fi.xyz = reinterpret_cast<const double*>(uintptr_t(0x20000));
                fi.cam = camera();
                fi.ro = ransac(cfg[1] == 200000 ? 16384 : 2048, f);
                fi.ro.progressive_sampling = f % 3 == 0;
                fi.ro.adaptive_stopping = f % 2;
                fi.bo = bundle(f % 5);
            }
            for (bool batched : {true, false}) {
                if (!batched && nf != 1) continue;
                Plan p;
                const Status s = p.build(in.data(), nf, tune, batched, 256, 2);
                expect(s.ok(), "plan build");
                if (!s.ok()) continue;
                std::vector<char> meta(p.meta_bytes() + 256);
                // the arena and the device metadata are addresses only here
                char* arena = reinterpret_cast<char*>(uintptr_t(1) << 40);
                char* meta_dev = reinterpret_cast<char*>(uintptr_t(1) << 41);
                p.bind(arena, meta_dev, meta.data());
            }
        }
        // a frame whose scoring work would need more than 2^26 items is
        // refused before the plan holds any of them
        {
            FrameInput fi;
            fi.n = (int64_t(1) << 28) - 1;
            fi.uv = reinterpret_cast<const double*>(uintptr_t(0x10000));
            fi.xyz = reinterpret_cast<const double*>(uintptr_t(0x20000));
            fi.cam = camera();
            fi.ro = ransac(1 << 24, 0);
            fi.bo = bundle(0);
            for (bool batched : {true, false}) {
                Plan big;
                expect(big.build(&fi, 1, tune, batched, 256, 2).code == TR_ERR_CAPACITY,
                       "2^26 scoring items");
            }
        }
        // one-frame plans rebuilt inside reserved capacity
        Plan p;
        p.reserve(Plan::single_counts(200000, 16384, true, tune));
        for (int64_t n : {3, 100, 5000, 200000}) {
            FrameInput fi;
            fi.n = n;
            fi.cam = camera();
            fi.ro = ransac(n == 200000 ? 16384 : 1024, 3);
            fi.ro.progressive_sampling = 1;
            fi.bo = bundle(TR_LOSS_HUBER);
            expect(p.fits(Plan::single_counts(n, fi.ro.max_iterations, true, tune)), "fits");
            expect(p.build(&fi, 1, tune, false, 256, 2).ok(), "single plan");
            std::vector<char> meta(p.meta_bytes() + 256);
            p.bind(reinterpret_cast<char*>(uintptr_t(1) << 40),
                   reinterpret_cast<char*>(uintptr_t(1) << 41), meta.data());
        }
    }
    return failures;
}
#endif

}  // namespace

int main(int argc, char** argv) {
    const std::string what = argc > 1 ? argv[1] : "";
    int f = -1;
#ifdef TR_CHECK_ORACLE
    if (what == "oracle") f = check_oracle();
#endif
#ifdef TR_CHECK_ENGINE
    if (what == "engine") f = check_engine();
#endif
    if (f < 0) {
        std::fprintf(stderr, "usage: host_checks oracle|engine (as built)\n");
        return 2;
    }
    std::printf("%s: %d failures\n", what.c_str(), f);
    return f == 0 ? 0 : 1;
}
