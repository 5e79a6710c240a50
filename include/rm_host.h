/*
 * rm_host.h — host-side I/O of the drop-in renderer CLI (librm_host.so, plain C ABI).
 *
 * No function here touches the GPU. The CLI `marcher.bin` (renderer_ray_marching_amd/host/)
 * is built from the same sources and calls the same code; this ABI exposes it to other
 * hosts and to the CPU tests.
 *
 * Status: 0 on success, a negative RMH_ERR_* code on failure; rmh_last_error() then holds
 * a message (per thread). Buffers returned through `**` arguments are malloc'ed by the
 * library and released with rmh_free.
 */
#ifndef RM_HOST_H
#define RM_HOST_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RMH_OK 0
#define RMH_ERR_INVALID (-1) /* bad argument or malformed input file */
#define RMH_ERR_IO (-2)      /* file could not be opened, read or written */

const char* rmh_last_error(void);
void rmh_free(void* p);

/* Load a PLY point cloud: vertex x, y, z (float or double properties) and red, green, blue
 * (uchar properties); ascii, binary_little_endian and binary_big_endian.
 * xyzw: n x 4 floats (x, y, z, 1); rgba: n x 4 floats (c / 255, 1).
 * Replaces Pointcloud::load_ply (src/pointcloud.cpp:43-58, happly getVertexPositions /
 * getVertexColors). */
int rmh_ply_load(const char* path, int64_t* n, float** xyzw, float** rgba);

/* Write a PLY (binary little-endian, float x y z, uchar red green blue). rgb: n x 3 bytes.
 * Used to produce inputs for the CLI; the reference has no writer. */
int rmh_ply_write(const char* path, const float* xyz, const uint8_t* rgb, int64_t n);

/* Radii cache <ply>.kdtree.radii: the text archive of a std::vector<float> that
 * main.cu:115-120 writes (boost::archive::text_oarchive) and main.cu:124-127 reads. The writer
 * emits "22 serialization::archive 17 0 0 <n> 0 v0 v1 ..." with every value as %.9e (exact
 * round trip); the reader accepts any library version (the item version is present above
 * version 3) and requires exactly <n> values after the header. */
int rmh_radii_write(const char* path, const float* radii, int64_t n);
int rmh_radii_read(const char* path, int64_t* n, float** radii);

/* Colour PNG: rgba is the H x W x 4 float texture (row 0 = bottom); written flipped to top-down
 * rows as RGBA8 with every channel (unsigned char)(v * 255.0f) and alpha forced to 255.
 * Replaces save_png (src/utils.cu:151-163). Values must lie in [0, 1]. */
int rmh_save_png(const char* path, const float* rgba, int W, int H);

/* Depth outputs of the same texture: depth = channel 3, flipped to top-down rows.
 * <path>: 8-bit grey PNG, 0 where |d - 1| < 1e-5, else (uint8)clamp(2.55f * d, 0, 255);
 * <path>.npy: float32 (H, W) array of the flipped depth.
 * Replaces save_depth (src/utils.cu:165-192). */
int rmh_save_depth(const char* path, const float* rgba, int W, int H);

/* ---- matrices JSON (main.cu:162-225) --------------------------------------------------- */

typedef struct rmh_matrices rmh_matrices;

/* One view entry: `key` is the target render path (the JSON key), split 0 = "train", 1 = "val".
 * Strings are owned by the rmh_matrices handle. */
typedef struct rmh_view_spec {
    const char* key;
    int32_t split;
    int32_t has_ply_path;   /* 1 if "source_scan_ply_path" is present */
    const char* ply_path;   /* its value, or NULL */
    float pose[16];         /* camera_pose, column-major (glm::from_json, src/utils.cu:137-149) */
    float calib[16];        /* calibration_mat, column-major */
    int32_t width, height;  /* (int)(2 * cx), (int)(2 * cy) (main.cu:202-203, 209) */
    float fov_radians;      /* 2 * atanf(width / (2 * fx)) (main.cu:204-206) */
} rmh_view_spec;

/* Parse the file. Both "train" and "val" objects are required (main.cu:220-225); views are
 * listed train first, then val, each in the sorted key order the reference iterates in
 * (nlohmann::json objects are std::map). */
int rmh_matrices_load(const char* path, rmh_matrices** out);
int64_t rmh_matrices_count(const rmh_matrices* m);
const char* rmh_matrices_key(const rmh_matrices* m, int64_t i);
/* Parse view i's parameters (done lazily, after the view's skip and lock checks, as in
 * main.cu:193-194): fails on a missing or malformed camera_pose / calibration_mat. */
int rmh_matrices_get(const rmh_matrices* m, int64_t i, rmh_view_spec* out);
void rmh_matrices_free(rmh_matrices* m);

/* ---- per-view output paths and locks (main.cu:172-192, 217-218) ------------------------ */

/* color = <output>/<parent dir of key>/<file of key>; depth = the same with every "_color" in
 * the file name replaced by "_depth"; lock = <output>/<parent>/.<file>.lock. Each buffer holds
 * `cap` bytes. */
int rmh_view_paths(const char* output_dir, const char* key, char* color, char* depth,
                   char* lock, size_t cap);

/* Create (truncate) the lock file and take an advisory whole-file write lock on it
 * (fcntl F_SETLK, what boost::interprocess::file_lock::try_lock does on POSIX). Returns 1 with
 * *fd set when the lock is taken, 0 when another process holds it, < 0 on error. */
int rmh_lock_try(const char* lock_path, int* fd);
/* Release the lock, close fd and remove the lock file. */
int rmh_lock_release(int fd, const char* lock_path);

/* ---- interactive viewer camera (no window) ------------------------------------------------ */
/* The render poses and fields of view of `frames` frames of the viewer's FPS camera
 * (src/camera.h:32-133, driven as main.cu:235-266 drives it) under a scripted input: one event
 * per line, "<frame> key W|S|A|D|Q|E", "<frame> mouse <dx> <dy>" or "<frame> scroll <dy>"
 * (script NULL: the built-in tour), with a fixed frame time dt (seconds). poses: frames x 16
 * column-major camera-to-world floats (inverse of the view matrix); fovs: frames radians. The
 * marcher.bin viewer renders exactly these poses. */
int rmh_camera_poses(const char* script, int frames, float dt, float* poses, float* fovs);

#ifdef __cplusplus
}
#endif
#endif
