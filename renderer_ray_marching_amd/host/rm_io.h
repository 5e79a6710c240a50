// Host-side I/O of the drop-in CLI: PLY, radii cache, matrices JSON, PNG / depth / .npy
// writers, per-view paths and locks. Plain C++17, no GPU code. Every function throws
// std::runtime_error (or std::filesystem::filesystem_error) on failure.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace rmh {

// ---- JSON ---------------------------------------------------------------------------------
// A small DOM. Objects are std::map, so they iterate in sorted key order, as nlohmann::json
// (the reference's parser, main.cu:166-167) does; a repeated key keeps its last value.
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object };
    Kind kind = Null;
    bool boolean = false;
    double number = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    const Json& at(const std::string& key) const;     // throws if absent or not an object
    const Json* find(const std::string& key) const;   // nullptr if absent
};
Json parse_json(const std::string& text);

// ---- PLY ----------------------------------------------------------------------------------
struct PlyCloud {
    int64_t n = 0;
    std::vector<float> xyzw;   // n x 4: x, y, z, 1
    std::vector<float> rgba;   // n x 4: r / 255, g / 255, b / 255, 1
};
PlyCloud load_ply(const std::string& path);
void write_ply(const std::string& path, const float* xyz, const uint8_t* rgb, int64_t n);

// ---- radii cache --------------------------------------------------------------------------
void write_radii(const std::string& path, const float* radii, int64_t n);
std::vector<float> read_radii(const std::string& path);

// ---- images -------------------------------------------------------------------------------
void save_png(const std::string& path, const float* rgba, int W, int H);
void save_depth(const std::string& path, const float* rgba, int W, int H);
// Building blocks, exposed for tests: an 8-bit PNG of `channels` (1 = grey, 4 = RGBA)
// top-down rows, and a float32 C-order .npy of shape (H, W).
void write_png8(const std::string& path, const uint8_t* pix, int W, int H, int channels);
void write_npy_f32(const std::string& path, const float* data, int H, int W);

// ---- matrices -----------------------------------------------------------------------------
struct ViewSpec {
    std::string key;
    int split = 0;                 // 0 = train, 1 = val
    bool has_ply_path = false;
    std::string ply_path;
    float pose[16];                // column-major
    float calib[16];               // column-major
    int width = 0, height = 0;
    float fov_radians = 0.f;
};
struct Matrices {
    Json root;
    std::vector<std::pair<int, const std::string*>> order;   // (split, key) in visiting order
    const Json& params(size_t i) const;
};
// Parse the file and list the views (throws if "train" or "val" is missing).
void load_matrices(const std::string& path, Matrices& m);
// Parse one view's parameters (throws on a missing or malformed matrix).
ViewSpec view_spec(const Matrices& m, size_t i);
// Row-major JSON 4x4 -> column-major floats (glm::from_json of src/utils.cu:137-149).
void matrix_from_json(const Json& j, float out[16]);

// ---- per-view paths and locks -------------------------------------------------------------
struct ViewPaths {
    std::string dir;     // <output>/<parent dir of key>
    std::string color;
    std::string depth;
    std::string lock;
};
ViewPaths view_paths(const std::string& output_dir, const std::string& key);
// 1 = taken (fd set), 0 = held by another process.
int lock_try(const std::string& lock_path, int* fd);
void lock_release(int fd, const std::string& lock_path);

// ---- radius clamp (main.cu:144-149) -------------------------------------------------------
inline float clamp_radius(float r, float max_radius) {
    return (r > max_radius) ? max_radius : r;
}

}  // namespace rmh
