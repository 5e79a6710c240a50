// The interactive viewer's camera and input, without a window (plain C++, no GPU code).
//
// FpsCamera restates the reference's FPS camera (src/camera.h:32-133): yaw / pitch Euler angles,
// Front = normalize(yaw_rot * pitch_rot * (0, 0, -1, 1)), Right = normalize(cross(Front, WorldUp)),
// Up = normalize(cross(Right, Front)), view = lookAt(Position, Position + Front, Up), and its
// keyboard, mouse and scroll handlers, quirks included (a keyboard step moves at least 0.1 m along
// every axis, because glm::max(vec3, 0.1f) is taken per component, src/camera.h:80-91; the D key
// moves LEFT and A moves RIGHT, main.cu:304-320). The render pose of a frame is inverse(view)
// and its field of view is Zoom (main.cu:249, 258).
//
// Input comes from a script instead of GLFW callbacks: one event per line,
//   <frame> key <W|S|A|D|Q|E>      the key is held during that frame (process_input, main.cu:304)
//   <frame> mouse <dx> <dy>        a left-button drag by (dx, dy) pixels (mouse_callback, :322)
//   <frame> scroll <dy>            a wheel step (scroll_callback, :340)
// '#' starts a comment. Events of a frame apply after that frame's render, keys first, in the
// reference's loop order (render, FPS count, process_input, then the callbacks of pollEvents).
#pragma once
#include <string>
#include <vector>

namespace rmh {

struct FpsCamera {
    float position[3] = {0.0f, 0.0f, 3.0f};   // main.cu:43
    float front[3] = {0.0f, 0.0f, -1.0f};
    float up[3] = {0.0f, 1.0f, 0.0f};
    float right[3] = {1.0f, 0.0f, 0.0f};
    float world_up[3] = {0.0f, 1.0f, 0.0f};
    float yaw = 0.0f, pitch = 0.0f;
    float speed = 2.5f;
    float sensitivity;        // 0.1 degree per pixel, in radians
    float zoom;               // field of view (radians), 60 degrees at the start (main.cu:249)

    FpsCamera();
    void update_vectors();
    // dir: 0 forward, 1 backward, 2 left, 3 right, 4 down, 5 up
    void keyboard(int dir, float dt);
    void mouse(float dx, float dy);
    void scroll(float dy);
    // Camera-to-world pose of the current frame (inverse of the lookAt view), column-major.
    void pose(float out[16]) const;
};

struct InputEvent {
    int frame = 0;
    int kind = 0;             // 0 key, 1 mouse, 2 scroll
    char key = 0;
    float a = 0.0f, b = 0.0f;
};

// Parse a script (throws std::runtime_error with the line number on a malformed line).
std::vector<InputEvent> parse_input_script(const std::string& text);
// The tour used without a script: walk forward while turning, look up and down, back away,
// strafe and rise, zoom in (120 frames).
std::vector<InputEvent> default_tour();
// Apply the events of one frame, in the reference's order.
void apply_frame_events(FpsCamera& cam, const std::vector<InputEvent>& ev, int frame, float dt);
// Key -> the reference's movement direction (W forward, S backward, D left, A right, Q down,
// E up); -1 for any other key.
int key_direction(char key);

// FPSCounter of src/utils.cu:104-112: counts frames and, once at least a second has passed since
// lastTime, prints "<count> FPS", advances lastTime by one second and restarts the count.
struct FpsCounter {
    double last_time = 0.0;
    unsigned counter = 0;
    // Returns true and sets *line when a line is due at time `now` (seconds since the start).
    bool update(double now, std::string* line);
};

}  // namespace rmh
