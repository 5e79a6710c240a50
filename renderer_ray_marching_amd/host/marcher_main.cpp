// marcher.bin — headless drop-in for the reference renderer's command line (main.cu:56-233).
//
// Same flags, files and per-view semantics: PLY in, <ply>.kdtree.radii cache (written unclamped,
// clamped after loading), matrices JSON ("train" then "val", sorted keys), per view
// <output>/<parent>/<file> colour PNG, "_color" -> "_depth" depth PNG plus <depth>.npy,
// skip-existing, advisory .<file>.lock files. The GPU work goes through librm_hip.so (include/rm.h):
// one tree build, one fused radii kernel, and the march: one view per launch by default, as the
// reference, through the batched kernel (packet traversal, warm start; the images are identical
// to the plain kernel's), -b N views per batched launch, or -b 0 for the plain per-view kernel.
//
// Deliberate differences (the reference's behaviour is undefined or fatal there):
//   * errors end the program with a message and exit status 1 instead of an uncaught exception;
//   * a cached radii file is remapped to a -s subsample only when a subsample was drawn, and
//     only if it holds one radius per point of the full PLY (main.cu:128-134 reads out of range
//     otherwise, and with -s equal to the point count sets every radius to the first one);
//   * matrices with fewer than 4 x 4 numbers are rejected (glm::from_json reads past them);
//   * a render key with a single path component writes into <output> itself;
//   * without -d there is no window to open: the viewer's compute half runs instead. Its FPS
//     camera (src/camera.h) is driven by a scripted input (--script, or a built-in tour), every
//     frame is rendered at 1024 x 768 (src/common.h:13-14) into a device buffer at
//     inverse(view) and the camera's zoom as field of view (main.cu:249-258) — by the batched
//     kernel through rm_render_batch_noalloc (nothing allocated per frame), or with -b 0 by the
//     plain kernel (rm_render), the same image either way — and "<n> FPS"
//     is printed once a second as FPSCounter::update does (src/utils.cu:104-112); --dump
//     writes frames and their poses for inspection.
#include <fcntl.h>
#include <hip/hip_runtime_api.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <iterator>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "rm.h"
#include "rm_io.h"
#include "viewer.h"

namespace fs = std::filesystem;
using std::cout;
using std::endl;
using std::string;

namespace {

// ---- command line (CLI11 conventions: -f X, -fX, --file X, --file=X, -dpi) ---------------

struct Args {
    string pcd_path, matrix_path, output_path;
    int max_points = -1;
    float max_radius = 0.1f;
    bool headless = false, precompute = false, ignore_existing = false;
    int batch = 1;
    bool plain = false;   // -b 0: the plain per-view kernel (rm_render)
    // viewer (without -d)
    string script_path, dump_dir;
    int frames = -1;      // -1: as many as the script covers (120 for the built-in tour)
    float dt = -1.0f;     // fixed frame time in seconds; < 0: the wall clock, as the reference
    int dump_every = 0;   // with --dump: every k-th frame (0: only the last)
};

// CLI11 exit codes for the errors this parser reports.
constexpr int kConversionError = 104;
constexpr int kExtrasError = 109;
constexpr int kArgumentMismatch = 114;

const char* kHelp =
    "Pointcloud Renderer\n"
    "Usage: marcher.bin [OPTIONS]\n"
    "\n"
    "Options:\n"
    "  -h,--help                   Print this help message and exit\n"
    "  -f,--file TEXT              Path to pointcloud to render\n"
    "  -m,--matrices TEXT          Path to view matrices json for which to render pointcloud in case of headless rendering.\n"
    "  -o,--output_path TEXT       Path where to store renders in case of headless rendering.\n"
    "  -s,--max_points INT         Take exact number of points.\n"
    "  -r,--max_radius FLOAT       Filter possible outliers in radii file by settings max radius.\n"
    "  -d,--headless               Run headlessly without a window\n"
    "  -p,--precompute-radii       Precompute radii even if already precomputed.\n"
    "  -i,--ignore_existing        Ignore existing renders and forcefully rewrite them.\n"
    "  -b,--batch INT              Views marched per batched GPU launch (default 1: one launch per view;\n"
    "                              0: the plain per-view kernel, also for the viewer's frames).\n"
    "Without -d (no window in this build: the viewer renders a scripted camera path):\n"
    "  --script TEXT               File of input events, one per line: '<frame> key W|S|A|D|Q|E',\n"
    "                              '<frame> mouse <dx> <dy>', '<frame> scroll <dy>' (default: a built-in tour).\n"
    "  --frames INT                Frames to render (default: to the script's last event).\n"
    "  --dt FLOAT                  Fixed frame time in seconds (default: the wall clock).\n"
    "  --dump TEXT                 Write frames (frame_NNNNN.npy, H x 4W float32, bottom-up) and\n"
    "                              their poses (frame_NNNNN.pose: 16 column-major floats, then the fov).\n"
    "  --dump-every INT            With --dump: every k-th frame (default: only the last).\n";

struct Opt {
    char short_name;
    const char* long_name;
    int kind;   // 0 = flag, 1 = string, 2 = int, 3 = float
    void* dst;
    bool seen = false;
};

// Returns -1 to continue, or the exit status.
int parse_args(int argc, char** argv, Args& a) {
    std::vector<Opt> opts = {
        {'f', "file", 1, &a.pcd_path},
        {'m', "matrices", 1, &a.matrix_path},
        {'o', "output_path", 1, &a.output_path},
        {'s', "max_points", 2, &a.max_points},
        {'r', "max_radius", 3, &a.max_radius},
        {'d', "headless", 0, &a.headless},
        {'p', "precompute-radii", 0, &a.precompute},
        {'i', "ignore_existing", 0, &a.ignore_existing},
        {'b', "batch", 2, &a.batch},
        {'\0', "script", 1, &a.script_path},
        {'\0', "frames", 2, &a.frames},
        {'\0', "dt", 3, &a.dt},
        {'\0', "dump", 1, &a.dump_dir},
        {'\0', "dump-every", 2, &a.dump_every},
    };
    auto label = [](const Opt& o) {
        return o.short_name ? string("-") + o.short_name + ",--" + o.long_name
                            : string("--") + o.long_name;
    };
    auto assign = [&](Opt& o, const string& v) -> int {
        if (o.kind != 0 && o.seen) {
            std::cerr << label(o) << ": at most 1 value allowed" << endl;
            return kArgumentMismatch;
        }
        o.seen = true;
        if (o.kind == 0) {
            bool val = true;
            if (v == "false" || v == "0" || v == "off" || v == "no") val = false;
            else if (!(v.empty() || v == "true" || v == "1" || v == "on" || v == "yes")) {
                std::cerr << "Could not convert: " << label(o) << " = " << v << endl;
                return kConversionError;
            }
            *(bool*)o.dst = val;
        } else if (o.kind == 1) {
            *(string*)o.dst = v;
        } else if (o.kind == 2) {
            char* end = nullptr;
            errno = 0;
            const long x = std::strtol(v.c_str(), &end, 10);
            if (v.empty() || *end || errno || x < INT32_MIN || x > INT32_MAX) {
                std::cerr << "Could not convert: " << label(o) << " = " << v << endl;
                return kConversionError;
            }
            *(int*)o.dst = (int)x;
        } else {
            char* end = nullptr;
            const float x = std::strtof(v.c_str(), &end);
            if (v.empty() || *end) {
                std::cerr << "Could not convert: " << label(o) << " = " << v << endl;
                return kConversionError;
            }
            *(float*)o.dst = x;
        }
        return -1;
    };
    std::vector<string> extras;
    for (int i = 1; i < argc; ++i) {
        const string s = argv[i];
        if (s == "-h" || s == "--help") {
            cout << kHelp;
            return 0;
        }
        if (s == "--") {
            for (++i; i < argc; ++i) extras.push_back(argv[i]);
            break;
        }
        if (s.size() > 2 && s[0] == '-' && s[1] == '-') {
            const size_t eq = s.find('=');
            const string name = s.substr(2, eq == string::npos ? string::npos : eq - 2);
            auto it = std::find_if(opts.begin(), opts.end(),
                                   [&](const Opt& o) { return name == o.long_name; });
            if (it == opts.end()) {
                extras.push_back(s);
                continue;
            }
            string v;
            if (eq != string::npos) {
                v = s.substr(eq + 1);
            } else if (it->kind != 0) {
                if (i + 1 >= argc) {
                    std::cerr << label(*it) << " requires 1 argument" << endl;
                    return kArgumentMismatch;
                }
                v = argv[++i];
            }
            if (int r = assign(*it, v); r >= 0) return r;
            continue;
        }
        if (s.size() > 1 && s[0] == '-') {
            for (size_t j = 1; j < s.size(); ++j) {
                auto it = std::find_if(opts.begin(), opts.end(), [&](const Opt& o) {
                    return o.short_name && s[j] == o.short_name;
                });
                if (it == opts.end()) {
                    extras.push_back(s);
                    break;
                }
                if (it->kind == 0) {
                    if (int r = assign(*it, ""); r >= 0) return r;
                    continue;
                }
                string v = s.substr(j + 1);
                if (v.empty()) {
                    if (i + 1 >= argc) {
                        std::cerr << label(*it) << " requires 1 argument" << endl;
                        return kArgumentMismatch;
                    }
                    v = argv[++i];
                }
                if (int r = assign(*it, v); r >= 0) return r;
                break;
            }
            continue;
        }
        extras.push_back(s);
    }
    if (!extras.empty()) {
        std::cerr << "The following arguments were not expected:";
        for (const string& e : extras) std::cerr << " " << e;
        std::cerr << endl << "Run with --help for more information." << endl;
        return kExtrasError;
    }
    if (a.batch < 0) {
        std::cerr << "-b,--batch must be at least 0" << endl;
        return kConversionError;
    }
    if (a.batch == 0) {
        a.plain = true;
        a.batch = 1;
    }
    if (a.frames < -1 || a.dump_every < 0) {
        std::cerr << "--frames and --dump-every must not be negative" << endl;
        return kConversionError;
    }
    return -1;
}

// ---- device plumbing ---------------------------------------------------------------------

struct Gpu {
    rm_ctx* ctx = nullptr;
    ~Gpu() {
        if (ctx) rm_ctx_destroy(ctx);
    }
    void check(int rc, const char* what) const {
        if (rc != RM_OK)
            throw std::runtime_error(string(what) + ": " + (ctx ? rm_last_error(ctx) : "no context"));
    }
};

struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t bytes) {
        if (p) (void)hipFree(p);
        p = nullptr;
        if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
            (void)hipGetLastError();
            throw std::runtime_error("device allocation of " + std::to_string(bytes) + " bytes failed");
        }
    }
    template <class T>
    T* as() const { return (T*)p; }
};

void to_device(void* dst, const void* src, size_t bytes) {
    if (bytes && hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("host-to-device copy failed");
}
void to_host(void* dst, const void* src, size_t bytes) {
    if (bytes && hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("device-to-host copy failed");
}

// ---- the view loop -----------------------------------------------------------------------

// An accepted view: locked, parameters parsed, waiting for its render.
struct Pending {
    rmh::ViewSpec spec;
    rmh::ViewPaths paths;
    int fd = -1;
};

// Releases a view's lock if it never completes (the lock file stays on disk, as when the
// reference's file_lock is destroyed during unwinding).
struct LockHolder {
    int fd = -1;
    LockHolder() = default;
    LockHolder(const LockHolder&) = delete;
    LockHolder& operator=(const LockHolder&) = delete;
    LockHolder(LockHolder&& o) noexcept : fd(o.fd) { o.fd = -1; }
    ~LockHolder() {
        if (fd >= 0) {
            struct flock fl;
            std::memset(&fl, 0, sizeof(fl));
            fl.l_type = F_UNLCK;
            fl.l_whence = SEEK_SET;
            ::fcntl(fd, F_SETLK, &fl);
            ::close(fd);
        }
    }
};

class ViewLoop {
public:
    ViewLoop(const Args& a, Gpu& gpu, rm_cloud* cloud) : a_(a), gpu_(gpu), cloud_(cloud) {}

    void run(const rmh::Matrices& m) {
        std::vector<Pending> group;
        std::vector<LockHolder> held;
        group.reserve(a_.batch);
        held.reserve(a_.batch);
        try {
            for (size_t i = 0; i < m.order.size(); ++i) {
                Pending p;
                if (!accept(m, i, p, held)) continue;
                group.push_back(std::move(p));
                if ((int)group.size() == a_.batch) {
                    render_group(group, held);
                    group.clear();
                    held.clear();
                }
            }
        } catch (...) {
            // views accepted before the failing one are rendered, then the loop ends
            if (!group.empty()) render_group(group, held);
            throw;
        }
        if (!group.empty()) render_group(group, held);
    }

private:
    // Skip and lock checks of main.cu:172-201. Returns true with the view locked and parsed.
    bool accept(const rmh::Matrices& m, size_t i, Pending& p, std::vector<LockHolder>& held) {
        const string& key = *m.order[i].second;
        p.paths = rmh::view_paths(a_.output_path, key);
        const fs::path output(a_.output_path);
        if (!fs::exists(output)) fs::create_directory(output);
        if (!fs::exists(p.paths.dir)) fs::create_directory(p.paths.dir);
        const fs::path color(p.paths.color);
        if (!a_.ignore_existing && fs::exists(color)) {
            cout << fs::canonical(fs::absolute(color)) << ": " << "ALREADY EXISTS" << endl;
            return false;
        }
        int fd = -1;
        if (!rmh::lock_try(p.paths.lock, &fd)) {
            cout << fs::absolute(color) << ": " << "ALREADY LOCKED" << endl;
            return false;
        }
        LockHolder guard;
        guard.fd = fd;
        cout << fs::absolute(color) << ": " << "LOCKING" << endl;
        p.spec = rmh::view_spec(m, i);
        const fs::path ply_for_view =
            fs::canonical(fs::absolute(fs::path(p.spec.has_ply_path ? p.spec.ply_path : a_.pcd_path)));
        const fs::path loaded = fs::canonical(fs::absolute(fs::path(a_.pcd_path)));
        if (ply_for_view != loaded) {
            cout << "Skipping " << loaded << ", rerun with proper ply." << endl;
            return false;
        }
        if (p.spec.width <= 0 || p.spec.height <= 0)
            throw std::runtime_error("view '" + key + "' has an empty image size");
        p.fd = fd;
        guard.fd = -1;
        held.emplace_back();
        held.back().fd = fd;
        return true;
    }

    void render_group(std::vector<Pending>& group, std::vector<LockHolder>& held) {
        std::vector<int64_t> offset(group.size() + 1, 0);
        for (size_t v = 0; v < group.size(); ++v)
            offset[v + 1] = offset[v] + (int64_t)group[v].spec.width * group[v].spec.height;
        const size_t bytes = sizeof(float) * 4 * (size_t)offset.back();
        if (rgba_bytes_ < bytes) {
            rgba_.alloc(bytes);
            rgba_bytes_ = bytes;
        }
        const auto start = std::chrono::high_resolution_clock::now();
        if (a_.plain) {
            const rmh::ViewSpec& s = group[0].spec;
            gpu_.check(rm_render(gpu_.ctx, cloud_, s.pose, s.fov_radians, s.width, s.height,
                                 rgba_.as<float>(), nullptr), "rm_render");
        } else {
            std::vector<rm_view> views(group.size());
            for (size_t v = 0; v < group.size(); ++v) {
                std::memcpy(views[v].pose, group[v].spec.pose, sizeof(views[v].pose));
                views[v].fov_radians = group[v].spec.fov_radians;
                views[v].width = group[v].spec.width;
                views[v].height = group[v].spec.height;
                views[v].reserved = 0;
                views[v].offset = offset[v];
            }
            gpu_.check(rm_render_batch(gpu_.ctx, cloud_, views.data(), (int)views.size(),
                                       rgba_.as<float>(), nullptr), "rm_render_batch");
        }
        gpu_.check(rm_ctx_synchronize(gpu_.ctx), "render");
        const auto end = std::chrono::high_resolution_clock::now();
        host_.resize(4 * (size_t)offset.back());
        to_host(host_.data(), rgba_.p, bytes);
        const float secs =
            (float)std::chrono::duration_cast<std::chrono::milliseconds>(end - start).count() / 1000.0f;
        if (group.size() > 1)
            cout << "batch of " << group.size() << " views: " << secs << " s" << endl;
        for (size_t v = 0; v < group.size(); ++v) {
            const Pending& p = group[v];
            const float* img = host_.data() + 4 * (size_t)offset[v];
            rmh::save_png(p.paths.color, img, p.spec.width, p.spec.height);
            rmh::save_depth(p.paths.depth, img, p.spec.width, p.spec.height);
            cout << fs::canonical(fs::absolute(fs::path(p.paths.color))) << ": "
                 << secs / (float)group.size() << " s" << endl;
            held[v].fd = -1;
            rmh::lock_release(p.fd, p.paths.lock);
        }
    }

    const Args& a_;
    Gpu& gpu_;
    rm_cloud* cloud_;
    DevBuf rgba_;
    size_t rgba_bytes_ = 0;
    std::vector<float> host_;
};

// ---- the viewer without a window ---------------------------------------------------------

// The render loop of main.cu:235-266 with scripted input: per frame, the frame time, the render
// of the camera's pose at 1024 x 768 (the reference's RGBA32F texture, here a device buffer),
// the FPS count, then the frame's input events.
void run_viewer(const Args& a, Gpu& gpu, rm_cloud* cloud) {
    constexpr int kW = 1024, kH = 768;   // SCREEN_WIDTH, SCREEN_HEIGHT (src/common.h:13-14)
    std::vector<rmh::InputEvent> events;
    if (!a.script_path.empty()) {
        std::ifstream in(a.script_path);
        if (!in) throw std::runtime_error("cannot read input script '" + a.script_path + "'");
        std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        events = rmh::parse_input_script(text);
    } else {
        events = rmh::default_tour();
    }
    int frames = a.frames;
    if (frames < 0) {
        frames = 0;
        for (const auto& e : events) frames = std::max(frames, e.frame + 1);
    }
    if (!a.dump_dir.empty()) fs::create_directories(a.dump_dir);
    DevBuf rgba((size_t)kW * kH * 16), hit((size_t)kW * kH * 4);
    std::vector<float> host;
    rmh::FpsCamera cam;
    rmh::FpsCounter fps;
    const auto t0 = std::chrono::steady_clock::now();
    auto seconds = [&] {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    float last_frame = 0.0f;
    for (int f = 0; f < frames; ++f) {
        const float current = (float)seconds();
        const float dt = a.dt >= 0.0f ? a.dt : current - last_frame;
        last_frame = current;
        float pose[16];
        cam.pose(pose);
        const float fov = cam.zoom;
        if (a.plain) {
            gpu.check(rm_render(gpu.ctx, cloud, pose, fov, kW, kH, rgba.as<float>(),
                                hit.as<int32_t>()), "rm_render");
        } else {
            rm_view v{};
            std::memcpy(v.pose, pose, sizeof(v.pose));
            v.fov_radians = fov;
            v.width = kW;
            v.height = kH;
            v.offset = 0;
            gpu.check(rm_render_batch_noalloc(gpu.ctx, cloud, &v, 1, rgba.as<float>(),
                                              hit.as<int32_t>()),
                      "rm_render_batch_noalloc");
        }
        gpu.check(rm_ctx_synchronize(gpu.ctx), "render");
        const bool dump = !a.dump_dir.empty() &&
                          (a.dump_every > 0 ? f % a.dump_every == 0 : f == frames - 1);
        if (dump) {
            host.resize((size_t)kW * kH * 4);
            to_host(host.data(), rgba.p, host.size() * sizeof(float));
            char name[32];
            std::snprintf(name, sizeof(name), "frame_%05d", f);
            const string base = (fs::path(a.dump_dir) / name).string();
            rmh::write_npy_f32(base + ".npy", host.data(), kH, 4 * kW);
            std::ofstream pf(base + ".pose");
            pf.precision(9);
            for (int i = 0; i < 16; ++i) pf << pose[i] << (i == 15 ? "\n" : " ");
            pf << fov << "\n";
            if (!pf) throw std::runtime_error("cannot write " + base + ".pose");
        }
        std::string line;
        if (fps.update(seconds(), &line)) cout << line << endl;
        rmh::apply_frame_events(cam, events, f, dt);
    }
    cout << "Rendered " << frames << " frames at " << kW << "x" << kH << " in " << seconds()
         << " s" << endl;
}

int run(const Args& a) {
    rmh::PlyCloud pc = rmh::load_ply(a.pcd_path);
    const int64_t n_full = pc.n;

    std::vector<int> max_points_indices((size_t)n_full);
    const bool subsampled = a.max_points > 0 && (int64_t)a.max_points < n_full;
    if (subsampled) {
        // main.cu:75-87: the same generator, seed and shuffle
        std::iota(max_points_indices.begin(), max_points_indices.end(), 0);
        std::mt19937 g(42);
        std::shuffle(max_points_indices.begin(), max_points_indices.end(), g);
        const int mp = a.max_points;
        std::vector<float> v((size_t)mp * 4), c((size_t)mp * 4);
        for (int ax = 0; ax < mp; ++ax) {
            std::memcpy(&v[4 * (size_t)ax], &pc.xyzw[4 * (size_t)max_points_indices[ax]], 16);
            std::memcpy(&c[4 * (size_t)ax], &pc.rgba[4 * (size_t)max_points_indices[ax]], 16);
        }
        pc.xyzw.swap(v);
        pc.rgba.swap(c);
        pc.n = mp;
    }
    const int64_t n = pc.n;
    cout << "Number of points: " << n << endl;

    Gpu gpu;
    gpu.check(rm_ctx_create(0, &gpu.ctx), "rm_ctx_create");
    DevBuf d_xyzw(16 * (size_t)n), d_rgba(16 * (size_t)n), d_r(4 * (size_t)n), d_rc(4 * (size_t)n);
    to_device(d_xyzw.p, pc.xyzw.data(), 16 * (size_t)n);
    to_device(d_rgba.p, pc.rgba.data(), 16 * (size_t)n);
    rm_tree* tree = nullptr;
    gpu.check(rm_tree_build(gpu.ctx, d_xyzw.as<float>(), n, 0, &tree), "rm_tree_build");
    struct TreeGuard {
        rm_tree* t;
        ~TreeGuard() { if (t) rm_tree_destroy(t); }
    } tree_guard{tree};

    const string radii_path = a.pcd_path + ".kdtree.radii";
    if (!fs::exists(fs::path(radii_path)) || a.precompute) {
        cout << "Precomputing radii..." << endl;
        gpu.check(rm_radii(gpu.ctx, tree, a.max_radius, d_r.as<float>(), d_rc.as<float>()), "rm_radii");
        gpu.check(rm_ctx_synchronize(gpu.ctx), "rm_radii");
        std::vector<float> radii_host((size_t)n);
        to_host(radii_host.data(), d_r.p, 4 * (size_t)n);
        rmh::write_radii(radii_path, radii_host.data(), n);   // unclamped, as main.cu:115-120
        if (!(a.max_radius > 0.0f)) to_device(d_rc.p, radii_host.data(), 4 * (size_t)n);
    } else {
        cout << "Using precomputed radii..." << endl;
        std::vector<float> radii_host = rmh::read_radii(radii_path);
        if (subsampled) {
            if ((int64_t)radii_host.size() != n_full)
                throw std::runtime_error("The radii file '" + radii_path + "' holds " +
                                         std::to_string(radii_host.size()) +
                                         " radii, not one per point of PLY file '" + a.pcd_path +
                                         "', so it cannot be subsampled (recompute it with -p).");
            std::vector<float> r_h((size_t)n);
            for (int64_t ax = 0; ax < n; ++ax) r_h[(size_t)ax] = radii_host[(size_t)max_points_indices[(size_t)ax]];
            radii_host.swap(r_h);
        }
        if ((int64_t)radii_host.size() != n)
            throw std::runtime_error(string("The radii file '") + radii_path +
                                     "' does not contain the same number of points as PLY file '" +
                                     a.pcd_path + "'.");
        if (a.max_radius > 0.0f)
            for (float& r : radii_host) r = rmh::clamp_radius(r, a.max_radius);
        to_device(d_rc.p, radii_host.data(), 4 * (size_t)n);
    }

    rm_cloud* cloud = nullptr;
    gpu.check(rm_cloud_create(gpu.ctx, tree, d_xyzw.as<float>(), d_rgba.as<float>(), d_rc.as<float>(), n,
                              &cloud), "rm_cloud_create");
    struct CloudGuard {
        rm_cloud* c;
        ~CloudGuard() { if (c) rm_cloud_destroy(c); }
    } cloud_guard{cloud};

    bool successful_run = true;
    try {
        if (a.headless) {
            if (!a.matrix_path.empty()) {
                std::ifstream probe(a.matrix_path);
                if (probe.good()) {
                    probe.close();
                    cout << "Matrices loaded." << endl;
                    rmh::Matrices m;
                    rmh::load_matrices(a.matrix_path, m);
                    ViewLoop loop(a, gpu, cloud);
                    loop.run(m);
                } else {
                    cout << "Error opening matrix file" << endl;
                }
            } else {
                cout << "The matrix file '" << a.matrix_path << "' was not found." << endl;
            }
        } else {
            run_viewer(a, gpu, cloud);
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << endl;
        successful_run = false;
    }
    return successful_run ? EXIT_SUCCESS : EXIT_FAILURE;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    if (int r = parse_args(argc, argv, a); r >= 0) return r;
    try {
        return run(a);
    } catch (const std::exception& e) {
        std::cerr << e.what() << endl;
        return EXIT_FAILURE;
    }
}
