// Matrices file, per-view output paths and advisory locks of the headless view loop
// (main.cu:162-225).
#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <regex>
#include <sstream>
#include <stdexcept>

#include "rm_io.h"

namespace fs = std::filesystem;

namespace rmh {

void matrix_from_json(const Json& j, float out[16]) {
    if (j.kind != Json::Array || j.arr.size() < 4)
        throw std::runtime_error("matrix must be an array of 4 rows");
    for (int r = 0; r < 4; ++r) {
        const Json& row = j.arr[r];
        if (row.kind != Json::Array || row.arr.size() < 4)
            throw std::runtime_error("matrix row " + std::to_string(r) + " must hold 4 numbers");
        for (int c = 0; c < 4; ++c) {
            if (row.arr[c].kind != Json::Number)
                throw std::runtime_error("matrix element is not a number");
            out[4 * c + r] = (float)row.arr[c].number;   // row-major JSON -> column-major
        }
    }
}

void load_matrices(const std::string& path, Matrices& m) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("could not open matrix file '" + path + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    m.root = parse_json(ss.str());
    m.order.clear();
    const char* splits[2] = {"train", "val"};
    for (int s = 0; s < 2; ++s) {
        const Json& part = m.root.at(splits[s]);
        if (part.kind != Json::Object)
            throw std::runtime_error(std::string("'") + splits[s] + "' must be an object");
        for (const auto& kv : part.obj) m.order.emplace_back(s, &kv.first);
    }
}

const Json& Matrices::params(size_t i) const {
    const auto& o = order.at(i);
    return root.at(o.first == 0 ? "train" : "val").at(*o.second);
}

ViewSpec view_spec(const Matrices& m, size_t i) {
    ViewSpec v;
    v.key = *m.order.at(i).second;
    v.split = m.order[i].first;
    const Json& p = m.params(i);
    matrix_from_json(p.at("camera_pose"), v.pose);
    matrix_from_json(p.at("calibration_mat"), v.calib);
    if (const Json* q = p.find("source_scan_ply_path")) {
        if (q->kind != Json::String)
            throw std::runtime_error("source_scan_ply_path must be a string");
        v.has_ply_path = true;
        v.ply_path = q->str;
    }
    // main.cu:202-206: W = 2 cx, H = 2 cy (glm [column][row] of the column-major matrix),
    // focal length fx only; the texture size truncates the float sizes.
    const float image_width = 2.0f * v.calib[4 * 2 + 0];
    const float image_height = 2.0f * v.calib[4 * 2 + 1];
    const float focal = v.calib[0];
    v.fov_radians = 2.0f * std::atan(image_width / (2.0f * focal));
    v.width = (int)image_width;
    v.height = (int)image_height;
    return v;
}

ViewPaths view_paths(const std::string& output_dir, const std::string& key) {
    const fs::path path(key);
    std::vector<fs::path> parts(path.begin(), path.end());
    if (parts.empty()) throw std::runtime_error("empty render path key");
    const fs::path last = parts.back();
    // A key of one component has no parent directory: the outputs go to <output> itself.
    const fs::path parent = parts.size() >= 2 ? parts[parts.size() - 2] : fs::path();
    const fs::path out(output_dir);
    ViewPaths v;
    v.dir = (out / parent).string();
    v.color = (out / parent / last).string();
    v.depth = (out / parent / std::regex_replace(last.string(), std::regex("_color"), "_depth")).string();
    v.lock = (out / parent / ("." + last.string() + ".lock")).string();
    return v;
}

int lock_try(const std::string& lock_path, int* fd_out) {
    { std::ofstream touch(lock_path); }   // create (or truncate) the lock file
    const int fd = ::open(lock_path.c_str(), O_RDWR);
    if (fd < 0) throw std::runtime_error("could not open lock file '" + lock_path + "': " + std::strerror(errno));
    struct flock fl;
    std::memset(&fl, 0, sizeof(fl));
    fl.l_type = F_WRLCK;
    fl.l_whence = SEEK_SET;
    fl.l_start = 0;
    fl.l_len = 0;   // whole file
    if (::fcntl(fd, F_SETLK, &fl) == -1) {
        const int e = errno;
        ::close(fd);
        if (e == EACCES || e == EAGAIN) return 0;
        throw std::runtime_error("could not lock '" + lock_path + "': " + std::strerror(e));
    }
    *fd_out = fd;
    return 1;
}

void lock_release(int fd, const std::string& lock_path) {
    struct flock fl;
    std::memset(&fl, 0, sizeof(fl));
    fl.l_type = F_UNLCK;
    fl.l_whence = SEEK_SET;
    ::fcntl(fd, F_SETLK, &fl);
    ::close(fd);
    std::error_code ec;
    fs::remove(lock_path, ec);
}

}  // namespace rmh
