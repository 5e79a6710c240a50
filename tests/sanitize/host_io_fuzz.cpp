// Parser robustness driver for the host drop-in library (librm_host sources), built by
// tests/test_host_sanitize.py with AddressSanitizer and UndefinedBehaviorSanitizer. For every
// input file it runs the matching parser on the file itself and on deterministic corruptions
// of it (every truncation length up to 512 bytes, then every 7th; 400 single-byte changes from
// a fixed-seed generator). A parser must either succeed or return an error code: a sanitizer
// report or a crash fails the test. It also writes PNG / depth / .npy outputs and exercises
// view path and lock helpers on valid and edge-case arguments.
//
// usage: host_io_fuzz <kind> <file> <scratch-dir>     kind: ply | radii | matrices | script | io
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <random>
#include <string>
#include <vector>

#include "rm_host.h"

namespace {

std::vector<char> slurp(const std::string& p) {
    std::ifstream f(p, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

void spit(const std::string& p, const std::vector<char>& b) {
    std::ofstream f(p, std::ios::binary | std::ios::trunc);
    f.write(b.data(), (std::streamsize)b.size());
}

// Runs the parser of `kind` on `path`; returns its status (the value itself is not checked).
int parse(const std::string& kind, const std::string& path) {
    if (kind == "ply") {
        int64_t n = 0;
        float *xyzw = nullptr, *rgba = nullptr;
        int r = rmh_ply_load(path.c_str(), &n, &xyzw, &rgba);
        if (r == RMH_OK) {
            volatile float s = 0.f;
            for (int64_t i = 0; i < 4 * n; ++i) s = s + xyzw[i] + rgba[i];
            rmh_free(xyzw);
            rmh_free(rgba);
        }
        return r;
    }
    if (kind == "radii") {
        int64_t n = 0;
        float* r = nullptr;
        int st = rmh_radii_read(path.c_str(), &n, &r);
        if (st == RMH_OK) {
            volatile float s = 0.f;
            for (int64_t i = 0; i < n; ++i) s = s + r[i];
            rmh_free(r);
        }
        return st;
    }
    if (kind == "matrices") {
        rmh_matrices* m = nullptr;
        int st = rmh_matrices_load(path.c_str(), &m);
        if (st == RMH_OK) {
            const int64_t c = rmh_matrices_count(m);
            for (int64_t i = 0; i < c; ++i) {
                rmh_view_spec v;
                (void)rmh_matrices_key(m, i);
                (void)rmh_matrices_get(m, i, &v);
            }
            (void)rmh_matrices_get(m, c, nullptr);
            rmh_matrices_free(m);
        }
        return st;
    }
    if (kind == "script") {
        std::vector<char> b = slurp(path);
        b.push_back('\0');
        float poses[16 * 8], fovs[8];
        return rmh_camera_poses(b.data(), 8, 0.016f, poses, fovs);
    }
    return -100;
}

int fuzz(const std::string& kind, const std::string& file, const std::string& dir) {
    const std::vector<char> orig = slurp(file);
    if (orig.empty()) {
        std::fprintf(stderr, "empty input %s\n", file.c_str());
        return 2;
    }
    if (parse(kind, file) != RMH_OK) {
        std::fprintf(stderr, "the valid input does not parse: %s\n", rmh_last_error());
        return 3;
    }
    const std::string tmp = dir + "/fuzz_input";
    int runs = 0;
    for (size_t len = 0; len < orig.size(); len += (len < 512 ? 1 : 7)) {
        spit(tmp, std::vector<char>(orig.begin(), orig.begin() + (std::ptrdiff_t)len));
        (void)parse(kind, tmp);
        ++runs;
    }
    std::mt19937 g(12345);
    for (int k = 0; k < 400; ++k) {
        std::vector<char> b = orig;
        const size_t at = g() % b.size();
        static const char vals[] = {'\0', '\n', ' ', '-', '9', 'e', '"', '{', ']', '\xff', '\x80'};
        b[at] = (k % 2) ? vals[g() % sizeof(vals)] : (char)(g() & 0xff);
        spit(tmp, b);
        (void)parse(kind, tmp);
        ++runs;
    }
    std::printf("%s: %d corrupted inputs parsed or rejected\n", kind.c_str(), runs);
    return 0;
}

int io(const std::string& dir) {
    // outputs of several sizes, including 1 x 1 and a non-multiple-of-4 width
    const int sizes[][2] = {{1, 1}, {3, 2}, {17, 5}, {64, 48}};
    for (const auto& s : sizes) {
        const int W = s[0], H = s[1];
        std::vector<float> rgba((size_t)W * H * 4);
        for (size_t i = 0; i < rgba.size(); ++i) rgba[i] = (float)((i * 37) % 101) / 50.0f - 0.3f;
        const std::string c = dir + "/c_" + std::to_string(W) + ".png";
        const std::string d = dir + "/d_" + std::to_string(W) + ".png";
        if (rmh_save_png(c.c_str(), rgba.data(), W, H) != RMH_OK) return 4;
        if (rmh_save_depth(d.c_str(), rgba.data(), W, H) != RMH_OK) return 5;
    }
    if (rmh_save_png((dir + "/x.png").c_str(), nullptr, 0, 0) == RMH_OK) return 6;
    char c[4096], d[4096], l[4096];
    const char* keys[] = {"a/b_color.png", "x", "/abs/p_color_color.jpg", "", "a/b/c/d_color.png"};
    for (const char* k : keys) (void)rmh_view_paths(dir.c_str(), k, c, d, l, sizeof c);
    (void)rmh_view_paths(dir.c_str(), "a/b_color.png", c, d, l, 4);   // too small a buffer
    int fd = -1;
    const std::string lk = dir + "/.lock";
    if (rmh_lock_try(lk.c_str(), &fd) == RMH_OK) (void)rmh_lock_release(fd, lk.c_str());
    // radii cache write/read round trip, including n = 0
    for (int64_t n : {0, 1, 5, 1000}) {
        std::vector<float> r((size_t)n);
        for (int64_t i = 0; i < n; ++i) r[(size_t)i] = 0.001f * (float)i;
        const std::string p = dir + "/r_" + std::to_string(n) + ".radii";
        if (rmh_radii_write(p.c_str(), r.data(), n) != RMH_OK) return 7;
        int64_t m = -1;
        float* back = nullptr;
        if (rmh_radii_read(p.c_str(), &m, &back) != RMH_OK || m != n) return 8;
        for (int64_t i = 0; i < n; ++i)
            if (back[i] != r[(size_t)i]) return 9;
        rmh_free(back);
    }
    std::printf("io: ok\n");
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s ply|radii|matrices|script|io <file> <scratch-dir>\n", argv[0]);
        return 1;
    }
    const std::string kind = argv[1];
    if (kind == "io") return io(argv[3]);
    return fuzz(kind, argv[2], argv[3]);
}
