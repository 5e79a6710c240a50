// The interactive viewer's camera and scripted input (see viewer.h). Vector algebra follows glm
// 0.9.9's definitions, evaluated left to right in float: normalize(v) = v * (1 / sqrt(dot(v, v))),
// rotate(I, angle, axis) from the axis-angle matrix, lookAt right-handed, and a general 4 x 4
// inverse (by the Laplace expansion in 2 x 2 minors). glm itself is not part of this checkout,
// so agreement with its last-bit rounding is not pinned.
#include "viewer.h"

#include <cmath>
#include <sstream>
#include <stdexcept>

namespace rmh {
namespace {

constexpr float kPi = 3.141592653589793238f;        // src/common.h:16
constexpr float kDegToRad = kPi / 180.0f;           // src/camera.h:26
constexpr float kZoomMax = kPi / 4;                 // src/camera.h:28

float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

void normalize3(const float* v, float* out) {
    const float s = 1.0f / std::sqrt(dot3(v, v));
    for (int i = 0; i < 3; ++i) out[i] = v[i] * s;
}

void cross3(const float* x, const float* y, float* out) {
    const float r0 = x[1] * y[2] - y[1] * x[2];
    const float r1 = x[2] * y[0] - y[2] * x[0];
    const float r2 = x[0] * y[1] - y[0] * x[1];
    out[0] = r0;
    out[1] = r1;
    out[2] = r2;
}

// rotate(mat4(1), angle, axis) for a unit axis: column c, row r at M[4c + r]
void rotation(float angle, const float* axis, float* M) {
    const float c = std::cos(angle), s = std::sin(angle);
    const float t[3] = {(1.0f - c) * axis[0], (1.0f - c) * axis[1], (1.0f - c) * axis[2]};
    const float R[3][3] = {
        {c + t[0] * axis[0], t[0] * axis[1] + s * axis[2], t[0] * axis[2] - s * axis[1]},
        {t[1] * axis[0] - s * axis[2], c + t[1] * axis[1], t[1] * axis[2] + s * axis[0]},
        {t[2] * axis[0] + s * axis[1], t[2] * axis[1] - s * axis[0], c + t[2] * axis[2]}};
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row)
            M[4 * col + row] = (col < 3 && row < 3) ? R[col][row] : (col == row ? 1.0f : 0.0f);
}

// column-major product A * B
void matmul4(const float* A, const float* B, float* out) {
    float r[16];
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row)
            r[4 * col + row] = A[row] * B[4 * col] + A[4 + row] * B[4 * col + 1] +
                               A[8 + row] * B[4 * col + 2] + A[12 + row] * B[4 * col + 3];
    for (int i = 0; i < 16; ++i) out[i] = r[i];
}

// Inverse of a column-major 4 x 4 (adjugate over the determinant, from the 2 x 2 minors of the
// first two and the last two rows). Returns false when the determinant is 0.
bool inverse4(const float* M, float* out) {
    float a[4][4];   // a[row][col]
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) a[r][c] = M[4 * c + r];
    const float s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
    const float s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
    const float s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
    const float s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
    const float s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
    const float s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
    const float c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
    const float c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
    const float c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
    const float c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
    const float c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
    const float c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
    const float det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    if (det == 0.0f) return false;
    const float inv = 1.0f / det;
    float b[4][4];
    b[0][0] = (a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * inv;
    b[0][1] = (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * inv;
    b[0][2] = (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * inv;
    b[0][3] = (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * inv;
    b[1][0] = (-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * inv;
    b[1][1] = (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * inv;
    b[1][2] = (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * inv;
    b[1][3] = (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * inv;
    b[2][0] = (a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * inv;
    b[2][1] = (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * inv;
    b[2][2] = (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * inv;
    b[2][3] = (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * inv;
    b[3][0] = (-a[1][0] * c3 + a[1][1] * c1 - a[1][2] * c0) * inv;
    b[3][1] = (a[0][0] * c3 - a[0][1] * c1 + a[0][2] * c0) * inv;
    b[3][2] = (-a[3][0] * s3 + a[3][1] * s1 - a[3][2] * s0) * inv;
    b[3][3] = (a[2][0] * s3 - a[2][1] * s1 + a[2][2] * s0) * inv;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) out[4 * c + r] = b[r][c];
    return true;
}

}  // namespace

FpsCamera::FpsCamera() : sensitivity(0.1f * kDegToRad), zoom(60.0f * (kPi / 180.0f)) {
    update_vectors();
}

void FpsCamera::update_vectors() {
    static const float ax[3] = {1.0f, 0.0f, 0.0f}, ay[3] = {0.0f, 1.0f, 0.0f};
    float P[16], Y[16], YP[16];
    rotation(pitch, ax, P);
    rotation(yaw, ay, Y);
    matmul4(Y, P, YP);
    const float look[4] = {0.0f, 0.0f, -1.0f, 1.0f};
    float f[3];
    for (int r = 0; r < 3; ++r)
        f[r] = YP[r] * look[0] + YP[4 + r] * look[1] + YP[8 + r] * look[2] + YP[12 + r] * look[3];
    normalize3(f, front);
    float rt[3];
    cross3(front, world_up, rt);
    normalize3(rt, right);
    float u[3];
    cross3(right, front, u);
    normalize3(u, up);
}

void FpsCamera::keyboard(int dir, float dt) {
    const float* axis = (dir <= 1) ? front : (dir <= 3) ? right : up;
    const float sign = (dir == 0 || dir == 3 || dir == 5) ? 1.0f : -1.0f;
    for (int i = 0; i < 3; ++i) {
        const float step = std::fmax(speed * axis[i] * dt, 0.1f);   // glm::max(vec3, float)
        position[i] = sign > 0 ? position[i] + step : position[i] - step;
    }
}

void FpsCamera::mouse(float dx, float dy) {
    yaw += dx * sensitivity;
    pitch += dy * sensitivity;
    const float lim = kPi / 2 - 0.1f;
    if (pitch > lim) pitch = lim;
    if (pitch < -lim) pitch = -lim;
    update_vectors();
}

void FpsCamera::scroll(float dy) {
    zoom -= dy * kDegToRad;
    if (zoom < 0.1f) zoom = 0.1f;
    if (zoom > kZoomMax) zoom = kZoomMax;
}

void FpsCamera::pose(float out[16]) const {
    // lookAt(eye, eye + front, up), right-handed
    float center[3], f[3], s[3], u[3], d[3];
    for (int i = 0; i < 3; ++i) center[i] = position[i] + front[i];
    for (int i = 0; i < 3; ++i) d[i] = center[i] - position[i];
    normalize3(d, f);
    cross3(f, up, d);
    normalize3(d, s);
    cross3(s, f, u);
    float V[16] = {s[0], u[0], -f[0], 0.0f, s[1], u[1], -f[1], 0.0f,
                   s[2], u[2], -f[2], 0.0f, -dot3(s, position), -dot3(u, position),
                   dot3(f, position), 1.0f};
    if (!inverse4(V, out)) throw std::runtime_error("camera view matrix is singular");
}

int key_direction(char key) {
    switch (key) {
        case 'W': return 0;
        case 'S': return 1;
        case 'D': return 2;
        case 'A': return 3;
        case 'Q': return 4;
        case 'E': return 5;
        default: return -1;
    }
}

std::vector<InputEvent> parse_input_script(const std::string& text) {
    std::vector<InputEvent> out;
    std::istringstream in(text);
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const auto hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ls(line);
        std::string kind;
        InputEvent e;
        if (!(ls >> e.frame)) {
            std::string rest;
            if (std::istringstream(line) >> rest)
                throw std::runtime_error("input script line " + std::to_string(lineno) +
                                         ": expected '<frame> key|mouse|scroll ...'");
            continue;   // blank or comment
        }
        bool ok = e.frame >= 0 && static_cast<bool>(ls >> kind);
        if (ok && kind == "key") {
            std::string k;
            ok = static_cast<bool>(ls >> k) && k.size() == 1 && key_direction(k[0]) >= 0;
            e.kind = 0;
            if (ok) e.key = k[0];
        } else if (ok && kind == "mouse") {
            e.kind = 1;
            ok = static_cast<bool>(ls >> e.a >> e.b);
        } else if (ok && kind == "scroll") {
            e.kind = 2;
            ok = static_cast<bool>(ls >> e.a);
        } else {
            ok = false;
        }
        std::string extra;
        if (!ok || (ls >> extra))
            throw std::runtime_error("input script line " + std::to_string(lineno) +
                                     ": expected '<frame> key W|S|A|D|Q|E', '<frame> mouse dx dy' "
                                     "or '<frame> scroll dy'");
        out.push_back(e);
    }
    return out;
}

std::vector<InputEvent> default_tour() {
    std::vector<InputEvent> ev;
    auto add = [&](int f, int kind, char key, float a, float b) {
        InputEvent e;
        e.frame = f;
        e.kind = kind;
        e.key = key;
        e.a = a;
        e.b = b;
        ev.push_back(e);
    };
    for (int f = 0; f < 40; ++f) {
        add(f, 0, 'W', 0, 0);
        add(f, 1, 0, 6.0f, 0.0f);
    }
    for (int f = 40; f < 60; ++f) add(f, 1, 0, 0.0f, f < 50 ? 8.0f : -8.0f);
    for (int f = 60; f < 80; ++f) add(f, 0, 'S', 0, 0);
    for (int f = 80; f < 100; ++f) {
        add(f, 0, 'A', 0, 0);
        add(f, 0, 'E', 0, 0);
    }
    for (int f = 100; f < 120; ++f) add(f, 2, 0, 1.0f, 0.0f);
    return ev;
}

void apply_frame_events(FpsCamera& cam, const std::vector<InputEvent>& ev, int frame, float dt) {
    for (const InputEvent& e : ev)
        if (e.frame == frame && e.kind == 0) cam.keyboard(key_direction(e.key), dt);
    for (const InputEvent& e : ev) {
        if (e.frame != frame) continue;
        if (e.kind == 1) cam.mouse(e.a, e.b);
        if (e.kind == 2) cam.scroll(e.a);
    }
}

bool FpsCounter::update(double now, std::string* line) {
    ++counter;
    if (now - last_time >= 1.0) {
        if (line) *line = std::to_string(counter) + " FPS";
        last_time += 1.0;
        counter = 0;
        return true;
    }
    return false;
}

}  // namespace rmh
