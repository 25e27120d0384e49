// hana_scene.cpp -- the interactive scene's host state: the orbit camera
// (Camera::update_transform, SR/camera.cpp:31-70), the key handlers of Scene
// and SingleModelScene (light rotation, shadow and shader switching,
// SR/scene.cpp:45-63,127-147) and the camera reset of update_camera
// (SR/camera.cpp:167-169).  Frame sequences driven by these feed
// hana_render_frame, batches or the server like the reference's main loop
// feeds DrawModel::draw.  Every float operation follows the reference's order
// (no contraction), so states match the reference's bit for bit.
#include <math.h>

#include <cmath>
#include <cstring>

#include "../../include/hana_gpu.h"

namespace {

struct V3 {
    float x, y, z;
};

// The reference's sin/cos/atan2/acos calls on floats resolve to the C
// library's float functions at run time; volatile pointers keep the compiler
// from evaluating them itself.
float (*volatile libm_sinf)(float) = ::sinf;
float (*volatile libm_cosf)(float) = ::cosf;
double (*volatile libm_tan)(double) = ::tan;
float (*volatile libm_atan2f)(float, float) = ::atan2f;
float (*volatile libm_acosf)(float) = ::acosf;
double (*volatile libm_pow)(double, double) = ::pow;

constexpr float kPi = 3.1415927f;        // SR/maths.h:7
constexpr float kEpsilon = 1e-5f;        // SR/maths.h:6
float radians(float deg) { return (kPi / 180) * deg; }   // TO_RADIANS, SR/maths.h:9

V3 v3(const float* p) { return {p[0], p[1], p[2]}; }
void store(V3 v, float* p) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 mul(V3 a, float f) { return {a.x * f, a.y * f, a.z * f}; }
// Vector::normal and normalize: sqrt((x*x + y*y) + z*z), v * (1 / normal)
// (SR/vector.h:41-42)
float norm(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
V3 normalize(V3 v) { return mul(v, 1.0f / norm(v)); }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float clampf(float f, float lo, float hi) { return f < lo ? lo : (f > hi ? hi : f); }   // SR/maths.cpp:3-5

const V3 kUp = {0, 1, 0};                 // SR/camera.cpp:14
const V3 kCameraPosition = {0, 0, 2.f};   // SR/camera.h:8
const V3 kCameraTarget = {0, 0, 0};       // SR/camera.h:9

// calculate_pan (SR/camera.cpp:31-42)
V3 pan_of(V3 from_camera, const hana_motion& m) {
    V3 forward = normalize(from_camera);
    V3 left = cross(kUp, forward);
    V3 up = cross(forward, left);
    float distance = norm(from_camera);
    // tan(FOVY / 2): FOVY is a constant of camera.cpp (SR/camera.cpp:13), so
    // the reference's build evaluates this call when compiling, correctly
    // rounded; glibc's tanf at run time is one ulp above that here.  The
    // double evaluation rounded to float gives the compile-time value.
    float half_fovy = radians(60) / 2;
    float factor = distance * (float)libm_tan((double)half_fovy) * 2;
    V3 dx = mul(mul(left, m.pan[0]), factor);
    V3 dy = mul(mul(up, m.pan[1]), factor);
    return add(dx, dy);
}

// calculate_offset (SR/camera.cpp:44-61)
V3 offset_of(V3 from_target, const hana_motion& m) {
    float radius = norm(from_target);
    float theta = libm_atan2f(from_target.x, from_target.z);   // azimuth
    float phi = libm_acosf(from_target.y / radius);           // polar
    float factor = kPi * 2;
    radius *= (float)libm_pow(0.95, (double)m.dolly);
    theta -= m.orbit[0] * factor;
    phi -= m.orbit[1] * factor;
    phi = clampf(phi, kEpsilon, kPi - kEpsilon);
    V3 o;
    o.x = radius * libm_sinf(phi) * libm_sinf(theta);
    o.y = radius * libm_cosf(phi);
    o.z = radius * libm_sinf(phi) * libm_cosf(theta);
    return o;
}

// rotate(angle, vx, vy, vz) (SR/maths.cpp:80-99) applied to (p, 1) and
// projected back to 3 components: out[i] = dot(row_i, (p, 1)) with the
// reverse-order dot of SR/vector.h:69-73.
V3 rotate_point(float angle, V3 axis, V3 p) {
    V3 n = normalize(axis);
    float c = libm_cosf(angle);
    float s = libm_sinf(angle);
    float m[3][3];
    m[0][0] = n.x * n.x * (1 - c) + c;
    m[0][1] = n.y * n.x * (1 - c) - s * n.z;
    m[0][2] = n.z * n.x * (1 - c) + s * n.y;
    m[1][0] = n.x * n.y * (1 - c) + s * n.z;
    m[1][1] = n.y * n.y * (1 - c) + c;
    m[1][2] = n.z * n.y * (1 - c) - s * n.x;
    m[2][0] = n.x * n.z * (1 - c) - s * n.y;
    m[2][1] = n.y * n.z * (1 - c) + s * n.x;
    m[2][2] = n.z * n.z * (1 - c) + c;
    float v[4] = {p.x, p.y, p.z, 1.f};
    float out[3];
    for (int i = 0; i < 3; i++) {
        float row[4] = {m[i][0], m[i][1], m[i][2], 0.0f};   // row i of the 4x4 (identity elsewhere)
        float r = 0.0f;
        for (int k = 4; k--;) r += row[k] * v[k];
        out[i] = r;
    }
    return {out[0], out[1], out[2]};
}

}  // namespace

extern "C" {

int hana_scene_init(hana_scene* s, int32_t width, int32_t height) {
    if (!s || width <= 0 || height <= 0) return HANA_E_INVALID;
    std::memset(s, 0, sizeof *s);
    // Scene::Scene (SR/scene.cpp:3-9) and SingleModelScene's first shader
    // (SR/scene.cpp:86,100)
    store(kCameraPosition, s->camera_position);
    store(kCameraTarget, s->camera_target);
    s->aspect = (float)width / (float)height;
    const float light[3] = {2, 2, 2};
    std::memcpy(s->light_position, light, sizeof light);
    s->enable_shadow = 1;
    s->shader_index = 0;
    return HANA_OK;
}

int hana_scene_update_camera(hana_scene* s, const hana_motion* m) {
    if (!s || !m) return HANA_E_INVALID;
    // Camera::update_transform (SR/camera.cpp:63-70)
    V3 position = v3(s->camera_position), target = v3(s->camera_target);
    V3 from_target = sub(position, target);
    V3 from_camera = sub(target, position);
    V3 pan = pan_of(from_camera, *m);
    V3 offset = offset_of(from_target, *m);
    target = add(target, pan);
    position = add(target, offset);
    store(position, s->camera_position);
    store(target, s->camera_target);
    return HANA_OK;
}

int hana_scene_key(hana_scene* s, int32_t key) {
    if (!s) return HANA_E_INVALID;
    switch (key) {
    case HANA_KEY_A:   // SR/scene.cpp:50-52
    case HANA_KEY_D:   // SR/scene.cpp:53-55
        store(rotate_point(radians(key == HANA_KEY_A ? -5 : 5), {0, 1, 1}, v3(s->light_position)), s->light_position);
        return HANA_OK;
    case HANA_KEY_E:   // SR/scene.cpp:56-58
        s->enable_shadow = !s->enable_shadow;
        return HANA_OK;
    case HANA_KEY_Q:   // SR/scene.cpp:134-139
        s->shader_index = (s->shader_index + 1 + 6) % 6;
        s->enable_shadow = s->shader_index < 3;
        return HANA_OK;
    case HANA_KEY_SPACE:   // update_camera's reset (SR/camera.cpp:167-169)
        store(kCameraPosition, s->camera_position);
        store(kCameraTarget, s->camera_target);
        return HANA_OK;
    default:
        return HANA_E_INVALID;
    }
}

int hana_scene_shader(const hana_scene* s) {
    // shaderInfos[0..5] (SR/scene.cpp:86-97)
    static const int kinds[6] = {HANA_SHADER_NORMALMAP, HANA_SHADER_BLINN, HANA_SHADER_TEXTURE_LIGHT,
                                 HANA_SHADER_TEXTURE, HANA_SHADER_TOON, HANA_SHADER_GROUND};
    if (!s || s->shader_index < 0 || s->shader_index > 5) return HANA_E_INVALID;
    return kinds[s->shader_index];
}

int hana_scene_uniforms(const hana_scene* s, const float obj_position[3], const float obj_rotation_deg[3],
                        const float obj_scale[3], int32_t fb_w, int32_t fb_h, hana_uniforms* out) {
    if (!s) return HANA_E_INVALID;
    return hana_build_uniforms(s->camera_position, s->camera_target, s->aspect, s->light_position, obj_position,
                               obj_rotation_deg, obj_scale, fb_w, fb_h, s->enable_shadow, out);
}

}  // extern "C"
