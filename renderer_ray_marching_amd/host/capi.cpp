// C ABI of the host I/O (declared in include/rm_host.h).
#include <cstdlib>
#include <cstring>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "rm_host.h"
#include "rm_io.h"
#include "viewer.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return RMH_OK;
    } catch (const std::bad_alloc&) {
        return fail(RMH_ERR_IO, "out of host memory");
    } catch (const std::exception& e) {
        return fail(RMH_ERR_INVALID, e.what());
    }
}

template <class T>
T* copy_out(const std::vector<T>& v) {
    T* p = (T*)std::malloc(v.empty() ? 1 : v.size() * sizeof(T));
    if (!p) throw std::bad_alloc();
    if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

}  // namespace

struct rmh_matrices {
    rmh::Matrices m;
    std::vector<std::string> ply_paths;   // keeps the strings handed out by rmh_matrices_get
};

extern "C" {

const char* rmh_last_error(void) { return g_err.c_str(); }

void rmh_free(void* p) { std::free(p); }

int rmh_ply_load(const char* path, int64_t* n, float** xyzw, float** rgba) {
    if (!path || !n || !xyzw || !rgba) return fail(RMH_ERR_INVALID, "rmh_ply_load: null argument");
    return guarded([&] {
        rmh::PlyCloud c = rmh::load_ply(path);
        float* a = copy_out(c.xyzw);
        float* b = nullptr;
        try {
            b = copy_out(c.rgba);
        } catch (...) {
            std::free(a);
            throw;
        }
        *n = c.n;
        *xyzw = a;
        *rgba = b;
    });
}

int rmh_ply_write(const char* path, const float* xyz, const uint8_t* rgb, int64_t n) {
    if (!path || (n > 0 && (!xyz || !rgb))) return fail(RMH_ERR_INVALID, "rmh_ply_write: null argument");
    return guarded([&] { rmh::write_ply(path, xyz, rgb, n); });
}

int rmh_radii_write(const char* path, const float* radii, int64_t n) {
    if (!path || (n > 0 && !radii)) return fail(RMH_ERR_INVALID, "rmh_radii_write: null argument");
    return guarded([&] { rmh::write_radii(path, radii, n); });
}

int rmh_radii_read(const char* path, int64_t* n, float** radii) {
    if (!path || !n || !radii) return fail(RMH_ERR_INVALID, "rmh_radii_read: null argument");
    return guarded([&] {
        std::vector<float> r = rmh::read_radii(path);
        *radii = copy_out(r);
        *n = (int64_t)r.size();
    });
}

int rmh_save_png(const char* path, const float* rgba, int W, int H) {
    if (!path || !rgba || W <= 0 || H <= 0) return fail(RMH_ERR_INVALID, "rmh_save_png: bad argument");
    return guarded([&] { rmh::save_png(path, rgba, W, H); });
}

int rmh_save_depth(const char* path, const float* rgba, int W, int H) {
    if (!path || !rgba || W <= 0 || H <= 0) return fail(RMH_ERR_INVALID, "rmh_save_depth: bad argument");
    return guarded([&] { rmh::save_depth(path, rgba, W, H); });
}

int rmh_matrices_load(const char* path, rmh_matrices** out) {
    if (!path || !out) return fail(RMH_ERR_INVALID, "rmh_matrices_load: null argument");
    return guarded([&] {
        rmh_matrices* h = new rmh_matrices();
        try {
            rmh::load_matrices(path, h->m);
        } catch (...) {
            delete h;
            throw;
        }
        h->ply_paths.resize(h->m.order.size());
        *out = h;
    });
}

int64_t rmh_matrices_count(const rmh_matrices* m) { return m ? (int64_t)m->m.order.size() : -1; }

const char* rmh_matrices_key(const rmh_matrices* m, int64_t i) {
    if (!m || i < 0 || i >= (int64_t)m->m.order.size()) return nullptr;
    return m->m.order[(size_t)i].second->c_str();
}

int rmh_matrices_get(const rmh_matrices* m, int64_t i, rmh_view_spec* out) {
    if (!m || !out || i < 0 || i >= (int64_t)m->m.order.size())
        return fail(RMH_ERR_INVALID, "rmh_matrices_get: bad argument");
    return guarded([&] {
        rmh::ViewSpec v = rmh::view_spec(m->m, (size_t)i);
        auto* mm = const_cast<rmh_matrices*>(m);
        mm->ply_paths[(size_t)i] = v.ply_path;
        out->key = m->m.order[(size_t)i].second->c_str();
        out->split = v.split;
        out->has_ply_path = v.has_ply_path ? 1 : 0;
        out->ply_path = v.has_ply_path ? mm->ply_paths[(size_t)i].c_str() : nullptr;
        std::memcpy(out->pose, v.pose, sizeof(v.pose));
        std::memcpy(out->calib, v.calib, sizeof(v.calib));
        out->width = v.width;
        out->height = v.height;
        out->fov_radians = v.fov_radians;
    });
}

void rmh_matrices_free(rmh_matrices* m) { delete m; }

int rmh_view_paths(const char* output_dir, const char* key, char* color, char* depth, char* lock,
                   size_t cap) {
    if (!output_dir || !key || !color || !depth || !lock || cap == 0)
        return fail(RMH_ERR_INVALID, "rmh_view_paths: bad argument");
    return guarded([&] {
        rmh::ViewPaths p = rmh::view_paths(output_dir, key);
        if (p.color.size() >= cap || p.depth.size() >= cap || p.lock.size() >= cap)
            throw std::runtime_error("rmh_view_paths: buffer too small");
        std::memcpy(color, p.color.c_str(), p.color.size() + 1);
        std::memcpy(depth, p.depth.c_str(), p.depth.size() + 1);
        std::memcpy(lock, p.lock.c_str(), p.lock.size() + 1);
    });
}

int rmh_lock_try(const char* lock_path, int* fd) {
    if (!lock_path || !fd) return fail(RMH_ERR_INVALID, "rmh_lock_try: null argument");
    int got = 0;
    const int rc = guarded([&] { got = rmh::lock_try(lock_path, fd); });
    if (rc != RMH_OK) return RMH_ERR_IO;
    return got;
}

int rmh_lock_release(int fd, const char* lock_path) {
    if (!lock_path || fd < 0) return fail(RMH_ERR_INVALID, "rmh_lock_release: bad argument");
    return guarded([&] { rmh::lock_release(fd, lock_path); });
}

int rmh_camera_poses(const char* script, int frames, float dt, float* poses, float* fovs) {
    if (frames < 0 || (frames > 0 && (!poses || !fovs)))
        return fail(RMH_ERR_INVALID, "rmh_camera_poses: bad arguments");
    return guarded([&] {
        const std::vector<rmh::InputEvent> ev =
            script ? rmh::parse_input_script(script) : rmh::default_tour();
        rmh::FpsCamera cam;
        for (int f = 0; f < frames; ++f) {
            cam.pose(poses + 16 * (size_t)f);
            fovs[f] = cam.zoom;
            rmh::apply_frame_events(cam, ev, f, dt);
        }
    });
}

}  // extern "C"
