// hana_hostmath.cpp -- see hana_hostmath.h.
#include "hana_hostmath.h"

#include <math.h>

#include <cmath>
#include <cstring>

#include "../../include/hana_gpu.h"

namespace hana {
namespace {

struct V3 { float x, y, z; };

// The reference's tan/cos/sin calls on floats resolve to the C library's
// tanf/cosf/sinf at run time.  Calling them through volatile pointers keeps
// the compiler from folding them at build time with its own (differently
// rounded) evaluation: glibc's tanf(0.5235988f) and a compile-time fold of
// the same call differ by one ulp.
float (*volatile libm_tanf)(float) = ::tanf;
float (*volatile libm_cosf)(float) = ::cosf;
float (*volatile libm_sinf)(float) = ::sinf;

// dot product, reverse index order from +0 (SR/vector.h:69-73)
float dot3(V3 a, V3 b) {
    float r = 0.0f;
    r += a.z * b.z;
    r += a.y * b.y;
    r += a.x * b.x;
    return r;
}
float dot4(const float* a, const float* b) {
    float r = 0.0f;
    for (int i = 4; i--;) r += a[i] * b[i];
    return r;
}
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// normalize: v * (1 / sqrt((x*x + y*y) + z*z)) (SR/vector.h:41-42)
V3 normalize(V3 v) {
    float s = 1.0f / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return {v.x * s, v.y * s, v.z * s};
}
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 v3(const float* p) { return {p[0], p[1], p[2]}; }

void identity(float* m) {
    for (int i = 0; i < 16; i++) m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
}

// determinant by cofactor expansion along row 0, cofactor sign applied as a
// float multiply by +-1 (SR/matrix.h:11-23, 72-87)
float det(const float* m, int n);
float cofactor(const float* m, int n, int row, int col) {
    float mn[9];
    int k = n - 1;
    for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) mn[i * k + j] = m[(i < row ? i : i + 1) * n + (j < col ? j : j + 1)];
    return det(mn, k) * (float)(((row + col) % 2) ? -1 : 1);
}
float det(const float* m, int n) {
    if (n == 1) return m[0];
    float r = 0.0f;
    for (int i = n; i--;) r += m[i] * cofactor(m, n, 0, i);
    return r;
}

constexpr float kPi = 3.1415927f;   // SR/maths.h:7
float radians(float deg) { return (kPi / 180) * deg; }   // TO_RADIANS, SR/maths.h:9

void rot_x(float a, float* m) { float c = libm_cosf(a), s = libm_sinf(a); identity(m); m[5] = c; m[6] = -s; m[9] = s; m[10] = c; }
void rot_y(float a, float* m) { float c = libm_cosf(a), s = libm_sinf(a); identity(m); m[0] = c; m[2] = s; m[8] = -s; m[10] = c; }
void rot_z(float a, float* m) { float c = libm_cosf(a), s = libm_sinf(a); identity(m); m[0] = c; m[1] = -s; m[4] = s; m[5] = c; }

}  // namespace

void m4_mul(const float* a, const float* b, float* out) {
    float r[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float col[4] = {b[j], b[4 + j], b[8 + j], b[12 + j]};
            r[4 * i + j] = dot4(a + 4 * i, col);
        }
    std::memcpy(out, r, sizeof r);
}

void m4_invert(const float* m, float* out) {
    float adj[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) adj[4 * i + j] = cofactor(m, 4, i, j);
    float d = dot4(adj, m);
    for (int i = 0; i < 16; i++) adj[i] = adj[i] / d;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = adj[4 * j + i];
}

void m4_lookat(const float* eye, const float* target, const float* up, float* m) {
    V3 e = v3(eye);
    V3 z = normalize(sub(e, v3(target)));
    V3 x = normalize(cross(v3(up), z));
    V3 y = cross(z, x);
    identity(m);
    m[0] = x.x; m[1] = x.y; m[2] = x.z;
    m[4] = y.x; m[5] = y.y; m[6] = y.z;
    m[8] = z.x; m[9] = z.y; m[10] = z.z;
    m[3] = -(dot3(x, e));
    m[7] = -(dot3(y, e));
    m[11] = -(dot3(z, e));
}

void m4_perspective(float fovy, float aspect, float n, float f, float* m) {
    float z_range = f - n;
    identity(m);
    m[5] = 1 / libm_tanf(fovy / 2);
    m[0] = m[5] / aspect;
    m[10] = -(n + f) / z_range;
    m[11] = -2 * n * f / z_range;
    m[14] = -1;
    m[15] = 0;
}

void m4_orthographic(float right, float top, float n, float f, float* m) {
    float z_range = f - n;
    identity(m);
    m[0] = 1 / right;
    m[5] = 1 / top;
    m[10] = -2 / z_range;
    m[11] = -(n + f) / z_range;
}

void m4_model(const float* pos, const float* rot_deg, const float* scale, float* out) {
    float t[16], rz[16], rx[16], ry[16], r[16], s[16], tmp[16];
    identity(t);
    t[3] = pos[0]; t[7] = pos[1]; t[11] = pos[2];
    rot_z(radians(rot_deg[2]), rz);
    rot_x(radians(rot_deg[0]), rx);
    rot_y(radians(rot_deg[1]), ry);
    m4_mul(rz, rx, tmp);
    m4_mul(tmp, ry, r);
    identity(s);
    s[0] = scale[0]; s[5] = scale[1]; s[10] = scale[2];
    m4_mul(t, r, tmp);
    m4_mul(tmp, s, out);
}

void build_uniforms(const float cam_pos[3], const float cam_target[3], float cam_aspect, const float light_pos[3],
                    const float obj_position[3], const float obj_rotation_deg[3], const float obj_scale[3], int fb_w,
                    int fb_h, int enable_shadow, hana_uniforms* u) {
    // Camera constants (SR/camera.cpp:11-14)
    const float kNear = 0.1f, kFar = 10000;
    const float kFovy = radians(60);
    const float up[3] = {0, 1, 0};
    const float origin[3] = {0, 0, 0};   // Camera::get_target_position (SR/camera.cpp:94-97)
    std::memset(u, 0, sizeof *u);
    m4_lookat(cam_pos, cam_target, up, u->view);
    m4_perspective(kFovy, cam_aspect, kNear, kFar, u->proj);
    m4_model(obj_position, obj_rotation_deg, obj_scale, u->model);
    m4_invert(u->model, u->model_inv);
    std::memcpy(u->view_pos, cam_pos, 3 * sizeof(float));
    V3 ld = normalize(sub(v3(light_pos), v3(origin)));
    u->light_dir[0] = ld.x; u->light_dir[1] = ld.y; u->light_dir[2] = ld.z;
    // LightColor / AMBIENT (SR/scene.h:8-9), material (SR/scene.cpp:81-84)
    const float lc[4] = {255.f / 255, 244.f / 255, 214.f / 255, 255};
    const float am[4] = {54.f / 255, 58.f / 255, 66.f / 255, 255};
    const float wh[4] = {1, 1, 1, 255};
    std::memcpy(u->light_color, lc, sizeof lc);
    std::memcpy(u->ambient, am, sizeof am);
    std::memcpy(u->mat_color, wh, sizeof wh);
    std::memcpy(u->mat_specular, wh, sizeof wh);
    u->gloss = 50;
    u->bump_scale = 1;
    float aspect = (float)fb_w / (float)fb_h;
    float ortho[16], look[16];
    m4_orthographic(aspect, 1, 0, 5, ortho);
    m4_lookat(light_pos, origin, up, look);
    m4_mul(ortho, look, u->light_vp);
    m4_mul(u->proj, u->view, u->camera_vp);
    u->enable_shadow = enable_shadow;
}

void derive_dev_uniforms(const hana_uniforms& u, DevUniforms& d) {
    std::memset(&d, 0, sizeof d);
    m4_mul(u.camera_vp, u.model, d.clip_m);
    m4_mul(u.light_vp, u.model, d.light_m);
    std::memcpy(d.model, u.model, sizeof d.model);
    std::memcpy(d.model_inv, u.model_inv, sizeof d.model_inv);
    std::memcpy(d.light_vp, u.light_vp, sizeof d.light_vp);
    std::memcpy(d.view_pos, u.view_pos, 3 * sizeof(float));
    std::memcpy(d.light_dir, u.light_dir, 3 * sizeof(float));
    std::memcpy(d.light_color, u.light_color, sizeof d.light_color);
    std::memcpy(d.ambient, u.ambient, sizeof d.ambient);
    std::memcpy(d.mat_color, u.mat_color, sizeof d.mat_color);
    std::memcpy(d.mat_spec, u.mat_specular, sizeof d.mat_spec);
    d.gloss = u.gloss;
    d.bump_scale = u.bump_scale;
    d.enable_shadow = u.enable_shadow;
    d.shadow_fp32 = u.shadow_fp32 ? 1 : 0;
}

}  // namespace hana

extern "C" int hana_build_uniforms(const float cam_pos[3], const float cam_target[3], float cam_aspect,
                                   const float light_pos[3], const float obj_position[3],
                                   const float obj_rotation_deg[3], const float obj_scale[3], int32_t fb_w,
                                   int32_t fb_h, int32_t enable_shadow, hana_uniforms* out) {
    if (!cam_pos || !cam_target || !light_pos || !obj_position || !obj_rotation_deg || !obj_scale || !out ||
        fb_w <= 0 || fb_h <= 0)
        return HANA_E_INVALID;
    hana::build_uniforms(cam_pos, cam_target, cam_aspect, light_pos, obj_position, obj_rotation_deg, obj_scale, fb_w,
                         fb_h, enable_shadow, out);
    return HANA_OK;
}
