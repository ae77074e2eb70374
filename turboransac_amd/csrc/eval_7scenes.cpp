// eval_7scenes.cpp - evaluator of scene files in the 7-Scenes correspondence
// format (TbrMain/src/main.cu:21-97, 160-202, 213-290; format README.md:54-62):
// a JSON object mapping frame names to {"gt_w2c": 4x4, "intrinsic K": 3x3,
// "point pair": [[u, v, X, Y, Z], ...]}.  Frames run in key order, as the
// reference's std::map does; each frame is estimated with the reference's
// options (1024 samples per round, success probability 0.9999, the rest at
// their defaults) and scored with its pose-error metric (main.cu:141-158).
// Per scene it prints the mean rotation error, the mean camera-centre error
// and the mean time per frame, and can write <scene>_pred.json with
// {"<frame>": {"pred_w2c": 4x4}}.
//
//   tr_eval7scenes [--data DIR] [--scenes a,b,...] [--batch] [--save-pred]
//                  [--max-iterations N] [--verbose] [--parse-only]
//
// --batch runs all frames of a scene in one tr_estimate_absolute_pose_batch
// call (per-frame time = total / frames); the default times one call per
// frame, transfers included, like the reference.  --parse-only reads the
// files and prints what it found without a GPU.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "turboransac.hpp"

namespace {

struct FrameData {
    double gt_w2c[16];
    double K[9];
    std::vector<double> p2d;  // N x 2
    std::vector<double> p3d;  // N x 3
};

// ---------------------------------------------------------------------------
// A pull reader for the scene files: numbers go straight into the frame's
// arrays, everything else is skipped.
// ---------------------------------------------------------------------------
class Reader {
  public:
    explicit Reader(const std::string& text) : s_(text), p_(0) {}
    bool ok() const { return err_.empty(); }
    const std::string& error() const { return err_; }

    void ws() {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\n' || s_[p_] == '\r' || s_[p_] == '\t'))
            ++p_;
    }
    bool peek(char c) {
        ws();
        return p_ < s_.size() && s_[p_] == c;
    }
    bool eat(char c) {
        if (peek(c)) {
            ++p_;
            return true;
        }
        return false;
    }
    bool expect(char c) {
        if (eat(c)) return true;
        fail(std::string("expected '") + c + "'");
        return false;
    }
    void fail(const std::string& m) {
        if (err_.empty()) err_ = m + " at byte " + std::to_string(p_);
    }
    bool is_null() {
        ws();
        if (s_.compare(p_, 4, "null") == 0) {
            p_ += 4;
            return true;
        }
        return false;
    }
    bool string(std::string* out) {
        if (!expect('"')) return false;
        std::string v;
        while (p_ < s_.size() && s_[p_] != '"') {
            char c = s_[p_++];
            if (c == '\\' && p_ < s_.size()) {
                const char e = s_[p_++];
                switch (e) {
                    case 'n': c = '\n'; break;
                    case 't': c = '\t'; break;
                    case 'r': c = '\r'; break;
                    case 'b': c = '\b'; break;
                    case 'f': c = '\f'; break;
                    case 'u': {  // keep the escape as written
                        v += "\\u";
                        continue;
                    }
                    default: c = e;
                }
            }
            v += c;
        }
        if (!expect('"')) return false;
        if (out) *out = v;
        return true;
    }
    bool number(double* out) {
        ws();
        const char* b = s_.c_str() + p_;
        char* e = nullptr;
        const double v = std::strtod(b, &e);
        if (e == b) {
            fail("expected a number");
            return false;
        }
        p_ += (size_t)(e - b);
        *out = v;
        return true;
    }
    // Skips any value.
    bool skip() {
        ws();
        if (p_ >= s_.size()) {
            fail("unexpected end");
            return false;
        }
        const char c = s_[p_];
        if (c == '"') return string(nullptr);
        if (c == '{' || c == '[') {
            const char close = c == '{' ? '}' : ']';
            ++p_;
            if (eat(close)) return true;
            do {
                if (c == '{') {
                    if (!string(nullptr) || !expect(':')) return false;
                }
                if (!skip()) return false;
            } while (eat(','));
            return expect(close);
        }
        for (const char* lit : {"true", "false", "null"}) {
            const size_t n = std::strlen(lit);
            if (s_.compare(p_, n, lit) == 0) {
                p_ += n;
                return true;
            }
        }
        double d;
        return number(&d);
    }
    // A rows x cols array of numbers into out (row-major).
    bool matrix(int rows, int cols, double* out) {
        if (!expect('[')) return false;
        for (int i = 0; i < rows; ++i) {
            if (i && !expect(',')) return false;
            if (!expect('[')) return false;
            for (int j = 0; j < cols; ++j) {
                if (j && !expect(',')) return false;
                if (!number(&out[i * cols + j])) return false;
            }
            if (!expect(']')) return false;
        }
        return expect(']');
    }
    // [[u, v, X, Y, Z], ...]; rows that are not five numbers are skipped
    // (main.cu:78-80).
    bool pairs(std::vector<double>* p2, std::vector<double>* p3) {
        if (!expect('[')) return false;
        if (eat(']')) return true;
        do {
            ws();
            if (!peek('[')) {
                if (!skip()) return false;
                continue;
            }
            ++p_;
            double v[5];
            int k = 0;
            bool numeric = true;
            if (!eat(']')) {
                do {
                    ws();
                    if (k < 5 && p_ < s_.size() &&
                        (s_[p_] == '-' || s_[p_] == '+' || s_[p_] == '.' || (s_[p_] >= '0' && s_[p_] <= '9'))) {
                        if (!number(&v[k])) return false;
                    } else {
                        numeric = false;
                        if (!skip()) return false;
                    }
                    ++k;
                } while (eat(','));
                if (!expect(']')) return false;
            }
            if (numeric && k == 5) {
                p2->push_back(v[0]);
                p2->push_back(v[1]);
                p3->push_back(v[2]);
                p3->push_back(v[3]);
                p3->push_back(v[4]);
            }
        } while (eat(','));
        return expect(']');
    }

  private:
    const std::string& s_;
    size_t p_;
    std::string err_;
};

// Frames with all three fields present and not null (main.cu:59-64).
bool load_scene(const std::string& path, std::map<std::string, FrameData>* frames,
                std::string* err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        *err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    Reader r(text);
    if (!r.expect('{')) {
        *err = r.error();
        return false;
    }
    if (!r.eat('}')) {
        do {
            std::string name;
            if (!r.string(&name) || !r.expect(':')) break;
            if (!r.peek('{')) {  // not a frame object
                if (!r.skip()) break;
                continue;
            }
            r.expect('{');
            FrameData fr;
            bool has_gt = false, has_k = false, has_pairs = false;
            if (!r.eat('}')) {
                do {
                    std::string key;
                    if (!r.string(&key) || !r.expect(':')) break;
                    if (r.is_null()) continue;
                    if (key == "gt_w2c") {
                        has_gt = r.matrix(4, 4, fr.gt_w2c);
                    } else if (key == "intrinsic K") {
                        has_k = r.matrix(3, 3, fr.K);
                    } else if (key == "point pair") {
                        has_pairs = r.pairs(&fr.p2d, &fr.p3d);
                    } else {
                        r.skip();
                    }
                } while (r.ok() && r.eat(','));
                r.expect('}');
            }
            if (!r.ok()) break;
            if (has_gt && has_k && has_pairs) (*frames)[name] = std::move(fr);
        } while (r.eat(','));
        r.expect('}');
    }
    if (!r.ok()) {
        *err = path + ": " + r.error();
        return false;
    }
    return true;
}

// Shepperd's method, normalised (the same conversion as the pose's).
void rotmat_to_quat(const double R[9], double q[4]) {
    const double tr = R[0] + R[4] + R[8];
    double w, x, y, z;
    if (tr > 0.0) {
        const double s = 2.0 * std::sqrt(tr + 1.0);
        w = 0.25 * s;
        x = (R[7] - R[5]) / s;
        y = (R[2] - R[6]) / s;
        z = (R[3] - R[1]) / s;
    } else if (R[0] > R[4] && R[0] > R[8]) {
        const double s = 2.0 * std::sqrt(1.0 + R[0] - R[4] - R[8]);
        w = (R[7] - R[5]) / s;
        x = 0.25 * s;
        y = (R[1] + R[3]) / s;
        z = (R[2] + R[6]) / s;
    } else if (R[4] > R[8]) {
        const double s = 2.0 * std::sqrt(1.0 + R[4] - R[0] - R[8]);
        w = (R[2] - R[6]) / s;
        x = (R[1] + R[3]) / s;
        y = 0.25 * s;
        z = (R[5] + R[7]) / s;
    } else {
        const double s = 2.0 * std::sqrt(1.0 + R[8] - R[0] - R[4]);
        w = (R[3] - R[1]) / s;
        x = (R[2] + R[6]) / s;
        y = (R[5] + R[7]) / s;
        z = 0.25 * s;
    }
    const double n = std::sqrt(w * w + x * x + y * y + z * z);
    q[0] = w / n;
    q[1] = x / n;
    q[2] = y / n;
    q[3] = z / n;
}

// Rotation error (degrees) of R_gt R_pred^T and camera-centre error (cm)
// between two world-to-camera 4x4 transforms (main.cu:141-158).
void pose_error(const double pred[16], const double gt[16], double* ae, double* te) {
    double Rp[9], Rg[9], E[9], Cp[3], Cg[3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Rp[3 * i + j] = pred[4 * i + j];
            Rg[3 * i + j] = gt[4 * i + j];
        }
    for (int i = 0; i < 3; ++i) {
        Cp[i] = -(Rp[i] * pred[3] + Rp[3 + i] * pred[7] + Rp[6 + i] * pred[11]);
        Cg[i] = -(Rg[i] * gt[3] + Rg[3 + i] * gt[7] + Rg[6 + i] * gt[11]);
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            E[3 * i + j] = Rg[3 * i] * Rp[3 * j] + Rg[3 * i + 1] * Rp[3 * j + 1] +
                           Rg[3 * i + 2] * Rp[3 * j + 2];
    double q[4];
    rotmat_to_quat(E, q);
    *ae = 2.0 * std::atan2(std::hypot(q[1], q[2], q[3]), std::fabs(q[0])) * 180.0 / M_PI;
    *te = std::sqrt((Cp[0] - Cg[0]) * (Cp[0] - Cg[0]) + (Cp[1] - Cg[1]) * (Cp[1] - Cg[1]) +
                    (Cp[2] - Cg[2]) * (Cp[2] - Cg[2])) *
          100.0;
}

void pose_to_w2c(const double q[4], const double t[3], double M[16]) {
    turboransac::CameraPose p;
    for (int i = 0; i < 4; ++i) p.q[i] = q[i];
    double R[9];
    p.R(R);
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) M[4 * i + j] = R[3 * i + j];
        M[4 * i + 3] = t[i];
    }
    M[12] = M[13] = M[14] = 0.0;
    M[15] = 1.0;
}

std::string json_escape(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

struct Result {
    std::string scene;
    double ae = 0, te = 0, ms = 0;
    int frames = 0, failed = 0;
};

}  // namespace

int main(int argc, char** argv) {
    std::string data_dir = "data/7scenes";
    std::vector<std::string> scenes = {"chess", "fire", "heads", "office",
                                       "pumpkin", "redkitchen", "stairs"};
    bool batch = false, save_pred = false, verbose = false, parse_only = false;
    uint64_t max_iterations = 1024;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char* what) -> std::string {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "%s needs a value\n", what);
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--data") data_dir = next("--data");
        else if (a == "--scenes") {
            scenes.clear();
            std::stringstream ss(next("--scenes"));
            std::string s;
            while (std::getline(ss, s, ','))
                if (!s.empty()) scenes.push_back(s);
        } else if (a == "--batch") batch = true;
        else if (a == "--save-pred") save_pred = true;
        else if (a == "--verbose") verbose = true;
        else if (a == "--parse-only") parse_only = true;
        else if (a == "--max-iterations") max_iterations = std::strtoull(next("--max-iterations").c_str(), nullptr, 10);
        else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }

    std::vector<Result> results;
    turboransac::Engine* engine = nullptr;
    for (const std::string& scene : scenes) {
        const std::string path = data_dir + "/" + scene + ".json";
        std::map<std::string, FrameData> frames;
        std::string err;
        if (!load_scene(path, &frames, &err)) {
            std::fprintf(stderr, "Failed to load JSON: %s\n", err.c_str());
            continue;
        }
        size_t pairs = 0;
        for (const auto& kv : frames) pairs += kv.second.p2d.size() / 2;
        std::printf("Loaded %zu valid frames (%zu correspondences) from %s\n", frames.size(), pairs,
                    path.c_str());
        if (parse_only || frames.empty()) continue;
        if (!engine) engine = new turboransac::Engine(0);

        // Options of the reference's evaluator (main.cu:177-182).
        turboransac::RansacOptions ro;
        ro.max_iterations = max_iterations;
        ro.success_prob = 0.9999;
        turboransac::BundleOptions bo;

        std::vector<std::string> names;
        std::vector<tr_frame> in;
        for (const auto& kv : frames) {
            const FrameData& fd = kv.second;
            tr_frame f{};
            f.points2D = fd.p2d.data();
            f.points3D = fd.p3d.data();
            f.num_points = (int64_t)(fd.p2d.size() / 2);
            f.camera.model_id = 1;
            f.camera.width = (int)fd.K[2] * 2;
            f.camera.height = (int)fd.K[5] * 2;
            f.camera.params[0] = fd.K[0];
            f.camera.params[1] = fd.K[4];
            f.camera.params[2] = fd.K[2];
            f.camera.params[3] = fd.K[5];
            f.ransac = ro;
            f.bundle = bo;
            names.push_back(kv.first);
            in.push_back(f);
        }
        const int nf = (int)in.size();
        std::vector<tr_pose_result> res(nf);
        std::vector<double> ms(nf, 0.0);
        std::vector<int> ok(nf, 1);
        if (batch) {
            const auto t0 = std::chrono::steady_clock::now();
            const tr_status s =
                tr_estimate_absolute_pose_batch(engine->handle(), in.data(), nf, res.data(), nullptr);
            const double total = std::chrono::duration<double, std::milli>(
                                     std::chrono::steady_clock::now() - t0).count();
            if (s != TR_OK) {
                std::fprintf(stderr, "%s: batch failed: %s (%s)\n", scene.c_str(), tr_status_string(s),
                             tr_context_last_error(engine->handle()));
                continue;
            }
            for (int i = 0; i < nf; ++i) {
                ms[i] = total / nf;
                ok[i] = res[i].status == TR_OK;
            }
        } else {
            for (int i = 0; i < nf; ++i) {
                const auto t0 = std::chrono::steady_clock::now();
                const tr_status s = tr_estimate_absolute_pose(
                    engine->handle(), in[i].points2D, in[i].points3D, in[i].num_points,
                    &in[i].camera, &in[i].ransac, &in[i].bundle, &res[i], nullptr);
                ms[i] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                            .count();
                ok[i] = s == TR_OK;
            }
        }

        Result R;
        R.scene = scene;
        std::string pred = "{\n";
        bool first = true;
        for (int i = 0; i < nf; ++i) {
            if (!ok[i]) {
                ++R.failed;
                continue;
            }
            double M[16], ae, te;
            pose_to_w2c(res[i].q, res[i].t, M);
            pose_error(M, frames[names[i]].gt_w2c, &ae, &te);
            R.ae += ae;
            R.te += te;
            R.ms += ms[i];
            ++R.frames;
            if (verbose)
                std::printf("%s: AE %.6f deg, TE %.6f cm, %.3f ms, inliers %llu\n", names[i].c_str(), ae,
                            te, ms[i], (unsigned long long)res[i].num_inliers);
            if (save_pred) {
                char buf[1024];
                pred += first ? "" : ",\n";
                first = false;
                pred += "  \"" + json_escape(names[i]) + "\": {\n    \"pred_w2c\": [";
                for (int r = 0; r < 4; ++r) {
                    std::snprintf(buf, sizeof buf, "%s[%.17g, %.17g, %.17g, %.17g]", r ? ", " : "",
                                  M[4 * r], M[4 * r + 1], M[4 * r + 2], M[4 * r + 3]);
                    pred += buf;
                }
                pred += "]\n  }";
            }
        }
        pred += "\n}\n";
        if (R.frames) {
            R.ae /= R.frames;
            R.te /= R.frames;
            R.ms /= R.frames;
        }
        results.push_back(R);
        if (save_pred) {
            const std::string out = data_dir + "/" + scene + "_pred.json";
            std::ofstream o(out);
            o << pred;
            std::printf("Saved predicted poses to %s\n", out.c_str());
        }
    }
    delete engine;

    if (!parse_only) {
        std::printf("\n=== Evaluation Results ===\n");
        for (const Result& r : results)
            std::printf("Scene: %-12s | Avg AE = %8.6f deg, Avg TE = %8.6f cm, Avg Time = %8.3f ms "
                        "(%4d frames, %d failed)\n",
                        r.scene.c_str(), r.ae, r.te, r.ms, r.frames, r.failed);
    }
    return 0;
}
