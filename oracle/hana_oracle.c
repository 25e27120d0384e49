/*
 * hana_oracle.c -- CPU restatement of the reference rasterisation path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the GPU product is compared against,
 * and the "port" CPU baseline.  It is never part of the product library.
 *
 * Every function keeps the reference's fp32 operation order; build with
 * -O2 -ffp-contract=off (see oracle/Makefile).  SR/x:N cites
 * Hana-SoftwareRenderer/x line N of the reference.
 *
 * Parity pins (tests/test_oracle.py): every entry point here is compared
 * bit for bit with the reference compiled from its own sources
 * (oracle/_ref, built by oracle/Makefile) and with the committed golden
 * vectors in tests/golden/ that the same compiled reference wrote.
 * powf is the C library's (glibc), as the reference's std::pow(float,float).
 */
#include "hana_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } v3;
typedef struct { float r, g, b, a; } rgba;

#define HO_EPSILON 1e-5f       /* SR/maths.h:6 */
#define HO_PI 3.1415927f       /* SR/maths.h:7 */

/* ------------------------------------------------------------------ */
/* Vector / Matrix arithmetic (SR/vector.h, SR/matrix.h)               */
/* ------------------------------------------------------------------ */

/* operator*(Vector,Vector): ret = 0; for (i = SIZE; i--;) ret += l[i]*r[i]
 * -- reverse index order (SR/vector.h:69-73). */
static float dot3(v3 a, v3 b) {
    float r = 0.0f;
    r += a.z * b.z;
    r += a.y * b.y;
    r += a.x * b.x;
    return r;
}
static float dot4(const float* a, const float* b) {
    float r = 0.0f;
    r += a[3] * b[3];
    r += a[2] * b[2];
    r += a[1] * b[1];
    r += a[0] * b[0];
    return r;
}
static float dot2(float ax, float ay, float bx, float by) {
    float r = 0.0f;
    r += ay * by;
    r += ax * bx;
    return r;
}

static v3 mk3(float x, float y, float z) { v3 v; v.x = x; v.y = y; v.z = z; return v; }
static v3 sub3(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }   /* SR/vector.h:64-67 */
static v3 add3(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }   /* SR/vector.h:59-62 */
static v3 scale3(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }     /* SR/vector.h:75-78 */

/* Vector<3>::normal: sqrt(x*x + y*y + z*z), forward order (SR/vector.h:41). */
static float norm3(v3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
/* Vector<3>::normalize(l = 1): *this = *this * (l / normal()) (SR/vector.h:42). */
static v3 normalize3(v3 v) { float s = 1.0f / norm3(v); return scale3(v, s); }
/* cross (SR/vector.h:97-99). */
static v3 cross3(v3 a, v3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* Matrix * Vector: ret[i] = row_i * v (SR/matrix.h:112-116). */
void ho_mat4_mul_vec4(const float* m, const float* v, float* out) {
    float r[4];
    for (int i = 4; i--;) r[i] = dot4(m + 4 * i, v);
    memcpy(out, r, sizeof r);
}
/* Matrix * Matrix: result[i][j] = lhs[i] * rhs.getCol(j) (SR/matrix.h:118-123). */
void ho_mat4_mul(const float* a, const float* b, float* out) {
    float r[16];
    for (int i = 4; i--;)
        for (int j = 4; j--;) {
            float col[4] = {b[j], b[4 + j], b[8 + j], b[12 + j]};
            r[4 * i + j] = dot4(a + 4 * i, col);
        }
    memcpy(out, r, sizeof r);
}

/* Minor / cofactor / determinant by expansion along row 0
 * (SR/matrix.h:11-23, 72-87): det = sum_{i=SIZE-1..0} m[0][i]*cofactor(0,i),
 * cofactor(r,c) = minor(r,c).det() * ((r+c)%2 ? -1 : 1). */
static float det_n(const float* m, int n);
static float cofactor_n(const float* m, int n, int row, int col) {
    float mn[9];
    int k = n - 1;
    for (int i = k; i--;)
        for (int j = k; j--;) mn[i * k + j] = m[(i < row ? i : i + 1) * n + (j < col ? j : j + 1)];
    return det_n(mn, k) * (float)(((row + col) % 2) ? -1 : 1);
}
static float det_n(const float* m, int n) {
    if (n == 1) return m[0];
    float ret = 0.0f;
    for (int i = n; i--;) ret += m[i] * cofactor_n(m, n, 0, i);
    return ret;
}
/* Matrix::invert = invert_transpose().transpose(); invert_transpose =
 * adjugate() / (adjugate()[0] * rows[0]) (SR/matrix.h:90-109). */
void ho_mat4_invert(const float* m, float* out) {
    float adj[16];
    for (int i = 4; i--;)
        for (int j = 4; j--;) adj[4 * i + j] = cofactor_n(m, 4, i, j);
    float tmp = dot4(adj, m);
    for (int i = 4; i--;)
        for (int k = 4; k--;) adj[4 * i + k] = adj[4 * i + k] / tmp;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = adj[4 * j + i];
}

static void identity4(float* m) {
    for (int i = 0; i < 16; i++) m[i] = (i % 5 == 0) ? 1.0f : 0.0f;
}

/* ------------------------------------------------------------------ */
/* Transform builders (SR/maths.cpp).  The reference's unqualified     */
/* cos/sin/tan calls take float arguments and resolve to the float     */
/* overloads, i.e. cosf/sinf/tanf.                                     */
/* ------------------------------------------------------------------ */

void ho_translate(float tx, float ty, float tz, float* m) {   /* SR/maths.cpp:39-45 */
    identity4(m);
    m[3] = tx;
    m[7] = ty;
    m[11] = tz;
}
void ho_scale(float sx, float sy, float sz, float* m) {       /* SR/maths.cpp:57-64 */
    identity4(m);
    m[0] = sx;
    m[5] = sy;
    m[10] = sz;
}
void ho_rotate(float angle, float vx, float vy, float vz, float* m) {  /* SR/maths.cpp:80-99 */
    v3 n = normalize3(mk3(vx, vy, vz));
    float c = cosf(angle);
    float s = sinf(angle);
    identity4(m);
    m[0] = n.x * n.x * (1 - c) + c;
    m[1] = n.y * n.x * (1 - c) - s * n.z;
    m[2] = n.z * n.x * (1 - c) + s * n.y;
    m[4] = n.x * n.y * (1 - c) + s * n.z;
    m[5] = n.y * n.y * (1 - c) + c;
    m[6] = n.z * n.y * (1 - c) - s * n.x;
    m[8] = n.x * n.z * (1 - c) - s * n.y;
    m[9] = n.y * n.z * (1 - c) + s * n.x;
    m[10] = n.z * n.z * (1 - c) + c;
}
void ho_rotate_x(float angle, float* m) {   /* SR/maths.cpp:111-120 */
    float c = cosf(angle), s = sinf(angle);
    identity4(m);
    m[5] = c;
    m[6] = -s;
    m[9] = s;
    m[10] = c;
}
void ho_rotate_y(float angle, float* m) {   /* SR/maths.cpp:132-141 */
    float c = cosf(angle), s = sinf(angle);
    identity4(m);
    m[0] = c;
    m[2] = s;
    m[8] = -s;
    m[10] = c;
}
void ho_rotate_z(float angle, float* m) {   /* SR/maths.cpp:153-162 */
    float c = cosf(angle), s = sinf(angle);
    identity4(m);
    m[0] = c;
    m[1] = -s;
    m[4] = s;
    m[5] = c;
}
void ho_lookat(const float* eye, const float* target, const float* up, float* m) {  /* SR/maths.cpp:180-195 */
    v3 e = mk3(eye[0], eye[1], eye[2]);
    v3 t = mk3(target[0], target[1], target[2]);
    v3 u = mk3(up[0], up[1], up[2]);
    v3 z = normalize3(sub3(e, t));
    v3 x = normalize3(cross3(u, z));
    v3 y = cross3(z, x);
    identity4(m);
    m[0] = x.x; m[1] = x.y; m[2] = x.z; m[3] = 0.0f;
    m[4] = y.x; m[5] = y.y; m[6] = y.z; m[7] = 0.0f;
    m[8] = z.x; m[9] = z.y; m[10] = z.z; m[11] = 0.0f;
    m[3] = -(dot3(x, e));
    m[7] = -(dot3(y, e));
    m[11] = -(dot3(z, e));
}
void ho_orthographic(float right, float top, float n, float f, float* m) {  /* SR/maths.cpp:214-223 */
    float z_range = f - n;
    identity4(m);
    m[0] = 1 / right;
    m[5] = 1 / top;
    m[10] = -2 / z_range;
    m[11] = -(n + f) / z_range;
}
void ho_perspective(float fovy, float aspect, float n, float f, float* m) {  /* SR/maths.cpp:243-255 */
    float z_range = f - n;
    identity4(m);
    m[5] = 1 / tanf(fovy / 2);
    m[0] = m[5] / aspect;
    m[10] = -(n + f) / z_range;
    m[11] = -2 * n * f / z_range;
    m[14] = -1;
    m[15] = 0;
}
/* viewport_transform (SR/maths.cpp:20-25). */
void ho_viewport_transform(int width, int height, const float* ndc, float* out) {
    out[0] = (ndc[0] + 1) * 0.5f * (float)width;
    out[1] = (ndc[1] + 1) * 0.5f * (float)height;
    out[2] = (ndc[2] + 1) * 0.5f;
}

#define TO_RADIANS(d) ((HO_PI / 180) * (d))   /* SR/maths.h:9 */

/* GameObject::GetModelMatrix = T * (Rz * Rx * Ry) * S (SR/gameobject.cpp:12-17). */
void ho_model_matrix(const float* p, const float* r, const float* s, float* out) {
    float t[16], rz[16], rx[16], ry[16], rot[16], sc[16], tmp[16];
    ho_translate(p[0], p[1], p[2], t);
    ho_rotate_z(TO_RADIANS(r[2]), rz);
    ho_rotate_x(TO_RADIANS(r[0]), rx);
    ho_rotate_y(TO_RADIANS(r[1]), ry);
    ho_mat4_mul(rz, rx, tmp);
    ho_mat4_mul(tmp, ry, rot);
    ho_scale(s[0], s[1], s[2], sc);
    ho_mat4_mul(t, rot, tmp);
    ho_mat4_mul(tmp, sc, out);
}

/* DrawModel::draw uniform set-up (SR/scene.h:55-71) with the camera
 * constants of SR/camera.cpp:11-14 and the colours of SR/scene.h:8-9 and
 * the material of SR/scene.cpp:81-84. */
void ho_build_uniforms(const float* cam_pos, const float* cam_target, float cam_aspect,
                       const float* light_pos, const float* obj_position,
                       const float* obj_rotation_deg, const float* obj_scale,
                       int fb_w, int fb_h, int enable_shadow, ho_uniforms* u) {
    static const float NEAR_ = 0.1f, FAR_ = 10000;
    const float FOVY = TO_RADIANS(60);
    const float up[3] = {0, 1, 0};
    const float zero[3] = {0, 0, 0};
    memset(u, 0, sizeof *u);
    ho_lookat(cam_pos, cam_target, up, u->view);
    ho_perspective(FOVY, cam_aspect, NEAR_, FAR_, u->proj);
    ho_model_matrix(obj_position, obj_rotation_deg, obj_scale, u->model);
    ho_mat4_invert(u->model, u->model_inv);
    memcpy(u->view_pos, cam_pos, 3 * sizeof(float));
    v3 ld = normalize3(sub3(mk3(light_pos[0], light_pos[1], light_pos[2]), mk3(zero[0], zero[1], zero[2])));
    u->light_dir[0] = ld.x; u->light_dir[1] = ld.y; u->light_dir[2] = ld.z;
    u->light_color[0] = 255.f / 255; u->light_color[1] = 244.f / 255;
    u->light_color[2] = 214.f / 255; u->light_color[3] = 255;
    u->ambient[0] = 54.f / 255; u->ambient[1] = 58.f / 255;
    u->ambient[2] = 66.f / 255; u->ambient[3] = 255;
    float aspect = (float)fb_w / (float)fb_h;
    float ortho[16], look[16];
    ho_orthographic(aspect, 1, 0, 5, ortho);
    ho_lookat(light_pos, zero, up, look);
    ho_mat4_mul(ortho, look, u->light_vp);
    ho_mat4_mul(u->proj, u->view, u->camera_vp);
    for (int i = 0; i < 3; i++) { u->mat_color[i] = 1.f; u->mat_specular[i] = 1.f; }
    u->mat_color[3] = 255; u->mat_specular[3] = 255;
    u->gloss = 50;
    u->bump_scale = 1;
    u->enable_shadow = enable_shadow;
}

/* ------------------------------------------------------------------ */
/* Color (SR/color.cpp) and textures                                    */
/* ------------------------------------------------------------------ */

static float smax(float a, float b) { return (a < b) ? b : a; }   /* std::max */
static float smin(float a, float b) { return (b < a) ? b : a; }   /* std::min */
static float saturate(float f) { return f < 0 ? 0 : (f > 1 ? 1 : f); }   /* SR/maths.cpp:7-9 */

static rgba mkc(float r, float g, float b, float a) { rgba c; c.r = r; c.g = g; c.b = b; c.a = a; return c; }
static rgba cadd(rgba x, rgba y) {   /* Color::operator+ (SR/color.cpp:38-45) */
    return mkc(smin(smax(0.f, x.r + y.r), 1.f), smin(smax(0.f, x.g + y.g), 1.f),
               smin(smax(0.f, x.b + y.b), 1.f), smin(smax(0.f, x.a + y.a), 1.f));
}
static rgba cmulf(rgba x, float i) {   /* Color::operator*(float) (SR/color.cpp:47-55) */
    i = (i > 1.f ? 1.f : (i < 0.f ? 0.f : i));
    return mkc(smin(smax(0.f, x.r * i), 1.f), smin(smax(0.f, x.g * i), 1.f),
               smin(smax(0.f, x.b * i), 1.f), smin(smax(0.f, x.a * i), 1.f));
}
static rgba cmulc(rgba x, rgba y) {   /* Color::operator*(Color) (SR/color.cpp:57-63) */
    return mkc(x.r * y.r, x.g * y.g, x.b * y.b, x.a * y.a);
}
static rgba c4(const float* p) { return mkc(p[0], p[1], p[2], p[3]); }

/* TGAImage::get + TGAColor(p, bpp): out-of-range or missing image -> all
 * zero (SR/tgaimage.cpp:248-253, SR/tgaimage.h:46-53). */
static void img_get(const ho_image* t, int x, int y, unsigned char bgra[4]) {
    bgra[0] = bgra[1] = bgra[2] = bgra[3] = 0;
    if (!t || !t->data || x < 0 || y < 0 || x >= t->width || y >= t->height) return;
    const unsigned char* p = t->data + (size_t)(x + y * t->width) * (size_t)t->bytespp;
    for (int i = 0; i < t->bytespp; i++) bgra[i] = p[i];
}
/* float -> int as the x86-64 conversion the reference compiles to
 * (cvttss2si): truncation, and INT_MIN for NaN / out of range. */
static int f2i(float f) {
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return (int)0x80000000u;
    return (int)f;
}
static int img_w(const ho_image* t) { return (t && t->data) ? t->width : 0; }
static int img_h(const ho_image* t) { return (t && t->data) ? t->height : 0; }

/* IShader::tex_diffuse -> Color(TGAColor) (SR/IShader.h:85-89, SR/color.cpp:5). */
static rgba tex_diffuse(const ho_image* t, const float* uv) {
    int x = f2i(uv[0] * img_w(t)), y = f2i(uv[1] * img_h(t));
    unsigned char c[4];
    img_get(t, x, y, c);
    return mkc(c[2] / 255.f, c[1] / 255.f, c[0] / 255.f, c[3] / 255.f);
}
/* IShader::tex_normal (SR/IShader.h:91-99): res[2-i] = c[i]/255*2-1. */
static v3 tex_normal(const ho_image* t, const float* uv) {
    int x = f2i(uv[0] * img_w(t)), y = f2i(uv[1] * img_h(t));
    unsigned char c[4];
    img_get(t, x, y, c);
    v3 r;
    r.z = (float)c[0] / 255.f * 2.f - 1.f;
    r.y = (float)c[1] / 255.f * 2.f - 1.f;
    r.x = (float)c[2] / 255.f * 2.f - 1.f;
    return r;
}

/* ------------------------------------------------------------------ */
/* Shaders (SR/IShader.h:49-130, SR/IShader.cpp)                        */
/* ------------------------------------------------------------------ */

typedef struct {
    const ho_uniforms* u;
    float clip_m[16];    /* camera_vp * model, ObjectToClipPos (SR/IShader.h:55-57) */
    float light_m[16];   /* light_vp * model, ObjectToViewPos (SR/IShader.h:59-61) */
    const ho_image* diffuse;
    const ho_image* normal;
    const ho_target* shadow_map;
} ctx_t;

static void ctx_init(ctx_t* c, const ho_uniforms* u, const ho_image* d, const ho_image* n, const ho_target* sm) {
    c->u = u;
    ho_mat4_mul(u->camera_vp, u->model, c->clip_m);
    ho_mat4_mul(u->light_vp, u->model, c->light_m);
    c->diffuse = d;
    c->normal = n;
    c->shadow_map = sm;
}

static void obj_to_clip(const float* m, const float* p, float* out) {
    float v[4] = {p[0], p[1], p[2], 1.0f};
    ho_mat4_mul_vec4(m, v, out);
}
static v3 obj_to_world_pos(const ho_uniforms* u, const float* p) {   /* SR/IShader.h:63-65 */
    float v[4] = {p[0], p[1], p[2], 1.0f}, o[4];
    ho_mat4_mul_vec4(u->model, v, o);
    return mk3(o[0], o[1], o[2]);
}
static v3 obj_to_world_normal(const ho_uniforms* u, const float* n) {   /* SR/IShader.h:71-75 */
    float row[4] = {n[0], n[1], n[2], 1.0f}, o[4];
    for (int j = 4; j--;) {
        float col[4] = {u->model_inv[j], u->model_inv[4 + j], u->model_inv[8 + j], u->model_inv[12 + j]};
        o[j] = dot4(row, col);
    }
    return mk3(o[0], o[1], o[2]);
}
static v3 light_dir(const ho_uniforms* u) { return mk3(u->light_dir[0], u->light_dir[1], u->light_dir[2]); }

static ho_v2f vertex_impl(const ctx_t* c, int kind, const ho_a2v* a) {
    ho_v2f o;
    memset(&o, 0, sizeof o);   /* Vector members zero-initialise; intensity is left 0 */
    const ho_uniforms* u = c->u;
    switch (kind) {
    case HO_SHADER_GROUND:
    case HO_SHADER_TOON: {   /* SR/IShader.cpp:5-9, 23-27 */
        obj_to_clip(c->clip_m, a->obj_pos, o.clip_pos);
        o.intensity = saturate(dot3(obj_to_world_normal(u, a->obj_normal), light_dir(u)));
        break;
    }
    case HO_SHADER_TEXTURE:   /* SR/IShader.cpp:47-52 */
        obj_to_clip(c->clip_m, a->obj_pos, o.clip_pos);
        o.uv[0] = a->uv[0]; o.uv[1] = a->uv[1];
        break;
    case HO_SHADER_TEXTURE_LIGHT: {   /* SR/IShader.cpp:65-71 */
        obj_to_clip(c->clip_m, a->obj_pos, o.clip_pos);
        v3 n = obj_to_world_normal(u, a->obj_normal);
        o.world_normal[0] = n.x; o.world_normal[1] = n.y; o.world_normal[2] = n.z;
        o.uv[0] = a->uv[0]; o.uv[1] = a->uv[1];
        break;
    }
    case HO_SHADER_BLINN:        /* SR/IShader.cpp:85-92 */
    case HO_SHADER_NORMALMAP: {  /* SR/IShader.cpp:117-124 */
        obj_to_clip(c->clip_m, a->obj_pos, o.clip_pos);
        v3 w = obj_to_world_pos(u, a->obj_pos);
        v3 n = obj_to_world_normal(u, a->obj_normal);
        o.world_pos[0] = w.x; o.world_pos[1] = w.y; o.world_pos[2] = w.z;
        o.world_normal[0] = n.x; o.world_normal[1] = n.y; o.world_normal[2] = n.z;
        o.uv[0] = a->uv[0]; o.uv[1] = a->uv[1];
        break;
    }
    case HO_SHADER_SHADOW:   /* SR/IShader.cpp:170-174 */
        obj_to_clip(c->light_m, a->obj_pos, o.clip_pos);
        break;
    default:
        break;
    }
    return o;
}

/* IShader::is_in_shadow (SR/IShader.h:107-129). */
static int is_in_shadow(const ctx_t* c, const float* dp, float n_dot_l) {
    const ho_target* sm = c->shadow_map;
    if (c->u->enable_shadow && sm) {
        float widht = (float)sm->width;
        float height = (float)sm->height;
        float ndc[3] = {dp[0] / dp[3], dp[1] / dp[3], dp[2] / dp[3]};
        float pos[3];
        ho_viewport_transform((int)widht, (int)height, ndc, pos);
        float depth_bias = 0.05f * (1 - n_dot_l);
        if (depth_bias < 0.005f) depth_bias = 0.01f;
        float current_depth = dp[2] - depth_bias;
        if (pos[0] < 0 || pos[1] < 0 || pos[0] >= widht || pos[1] >= height) return 1;
        int px = (int)pos[0], py = (int)pos[1];
        float closest_depth;
        if (c->u->shadow_fp32)   /* extension: the nearest caster's light clip z from the fp32 depth */
            closest_depth = 2.f * sm->depth[(size_t)py * sm->width + px] - 1.f;
        else
            closest_depth = sm->color[((size_t)py * sm->width + px) * 4 + 0] / 255.f;
        return current_depth < closest_depth;
    }
    return 1;
}

/* Shared tail of BlinnShader/NormalMapShader::fragment (SR/IShader.cpp:97-108, 150-161). */
static rgba blinn_tail(const ctx_t* c, const ho_v2f* v, v3 n) {
    const ho_uniforms* u = c->u;
    rgba albedo = cmulc(tex_diffuse(c->diffuse, v->uv), c4(u->mat_color));
    rgba ambient = cmulc(c4(u->ambient), albedo);
    float n_dot_l = saturate(dot3(n, light_dir(u)));
    rgba diffuse = cmulf(cmulc(c4(u->light_color), albedo), n_dot_l);
    v3 view_dir = normalize3(sub3(mk3(u->view_pos[0], u->view_pos[1], u->view_pos[2]),
                                  mk3(v->world_pos[0], v->world_pos[1], v->world_pos[2])));
    v3 half_dir = normalize3(add3(view_dir, light_dir(u)));
    rgba spec = cmulf(cmulc(c4(u->light_color), c4(u->mat_specular)),
                      powf(saturate(dot3(n, half_dir)), u->gloss));
    float wp[4] = {v->world_pos[0], v->world_pos[1], v->world_pos[2], 1.0f}, dp[4];
    ho_mat4_mul_vec4(u->light_vp, wp, dp);
    int shadow = is_in_shadow(c, dp, n_dot_l);
    return cadd(ambient, cmulf(cadd(diffuse, spec), (float)shadow));
}

static rgba fragment_impl(const ctx_t* c, int kind, ho_v2f* v) {
    const ho_uniforms* u = c->u;
    static const rgba WHITE = {1.f, 1.f, 1.f, 255.f};
    switch (kind) {
    case HO_SHADER_GROUND:   /* SR/IShader.cpp:11-15 */
        return cmulf(WHITE, v->intensity);
    case HO_SHADER_TOON: {   /* SR/IShader.cpp:29-39 */
        float intensity = v->intensity;
        if (intensity > .85) intensity = 1;
        else if (intensity > .60) intensity = (float).80;
        else if (intensity > .45) intensity = (float).60;
        else if (intensity > .30) intensity = (float).45;
        else if (intensity > .15) intensity = (float).30;
        return cmulf(mkc(1, 155 / 255.f, 0, 255), intensity);
    }
    case HO_SHADER_TEXTURE:   /* SR/IShader.cpp:53-57 */
        return tex_diffuse(c->diffuse, v->uv);
    case HO_SHADER_TEXTURE_LIGHT: {   /* SR/IShader.cpp:72-77 */
        float intensity = saturate(dot3(mk3(v->world_normal[0], v->world_normal[1], v->world_normal[2]), light_dir(u)));
        return cmulf(tex_diffuse(c->diffuse, v->uv), intensity);
    }
    case HO_SHADER_BLINN: {   /* SR/IShader.cpp:94-109 */
        v3 n = normalize3(mk3(v->world_normal[0], v->world_normal[1], v->world_normal[2]));
        v->world_normal[0] = n.x; v->world_normal[1] = n.y; v->world_normal[2] = n.z;
        return blinn_tail(c, v, n);
    }
    case HO_SHADER_NORMALMAP: {   /* SR/IShader.cpp:126-162 */
        v3 normal = mk3(v->world_normal[0], v->world_normal[1], v->world_normal[2]);
        float x = normal.x, y = normal.y, z = normal.z;
        v3 t = mk3(x * y / sqrtf(x * x + z * z), sqrtf(x * x + z * z), z * y / sqrtf(x * x + z * z));
        v3 b = cross3(normal, t);
        v3 r0 = mk3(t.x, b.x, normal.x), r1 = mk3(t.y, b.y, normal.y), r2 = mk3(t.z, b.z, normal.z);
        v3 bump = tex_normal(c->normal, v->uv);
        bump.x = bump.x * u->bump_scale;
        bump.y = bump.y * u->bump_scale;
        bump.z = (float)sqrt(1.0 - saturate(dot2(bump.x, bump.y, bump.x, bump.y)));
        v3 tb;
        tb.z = dot3(r2, bump);
        tb.y = dot3(r1, bump);
        tb.x = dot3(r0, bump);
        return blinn_tail(c, v, normalize3(tb));
    }
    case HO_SHADER_SHADOW: {   /* SR/IShader.cpp:176-180 */
        float factor = v->clip_pos[2];
        return cmulf(WHITE, factor);
    }
    default:
        return mkc(0, 0, 0, 0);
    }
}

ho_v2f ho_vertex(int kind, const ho_uniforms* u, const ho_a2v* a2v) {
    ctx_t c;
    ctx_init(&c, u, NULL, NULL, NULL);
    return vertex_impl(&c, kind, a2v);
}

int ho_fragment(int kind, const ho_uniforms* u, const ho_image* diffuse, const ho_image* normal,
                const ho_target* shadow_map, const ho_v2f* v2f, float* out) {
    ctx_t c;
    ctx_init(&c, u, diffuse, normal, shadow_map);
    ho_v2f v = *v2f;
    rgba r = fragment_impl(&c, kind, &v);
    out[0] = r.r; out[1] = r.g; out[2] = r.b; out[3] = r.a;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Pipeline (SR/graphics.cpp)                                          */
/* ------------------------------------------------------------------ */

enum { PW, PX, NX, PY, NY, PZ, NZ };

static int is_inside_plane(const float* c, int plane) {   /* SR/graphics.cpp:29-49 */
    switch (plane) {
    case PW: return c[3] >= HO_EPSILON;
    case PX: return c[0] <= +c[3];
    case NX: return c[0] >= -c[3];
    case PY: return c[1] <= +c[3];
    case NY: return c[1] >= -c[3];
    case PZ: return c[2] <= +c[3];
    default: return c[2] >= -c[3];
    }
}
static float intersect_ratio(const float* p, const float* c, int plane) {   /* SR/graphics.cpp:51-71 */
    switch (plane) {
    case PW: return (p[3] - HO_EPSILON) / (p[3] - c[3]);
    case PX: return (p[3] - p[0]) / ((p[3] - p[0]) - (c[3] - c[0]));
    case NX: return (p[3] + p[0]) / ((p[3] + p[0]) - (c[3] + c[0]));
    case PY: return (p[3] - p[1]) / ((p[3] - p[1]) - (c[3] - c[1]));
    case NY: return (p[3] + p[1]) / ((p[3] + p[1]) - (c[3] + c[1]));
    case PZ: return (p[3] - p[2]) / ((p[3] - p[2]) - (c[3] - c[2]));
    default: return (p[3] + p[2]) / ((p[3] + p[2]) - (c[3] + c[2]));
    }
}
/* clip_against_plane + liner_interpolate_varyings (SR/graphics.cpp:7-17, 73-110). */
static int clip_against_plane(int plane, int n, const ho_v2f* in, ho_v2f* out) {
    int m = 0;
    for (int i = 0; i < n; i++) {
        int prev = (i - 1 + n) % n;
        const float* pc = in[prev].clip_pos;
        const float* cc = in[i].clip_pos;
        int pin = is_inside_plane(pc, plane), cin = is_inside_plane(cc, plane);
        if (pin != cin) {
            float t = intersect_ratio(pc, cc, plane);
            const float* f = (const float*)&in[prev];
            const float* to = (const float*)&in[i];
            float* d = (float*)&out[m];
            for (int k = 0; k < 13; k++) d[k] = f[k] + (to[k] - f[k]) * t;
            m++;
        }
        if (cin) out[m++] = in[i];
    }
    return m;
}
static int is_vertex_visible(const float* v) {   /* SR/graphics.cpp:132-134 */
    return fabsf(v[0]) <= v[3] && fabsf(v[1]) <= v[3] && fabsf(v[2]) <= v[3];
}
/* clip_triangle (SR/graphics.cpp:136-161) with in/out buffers of 30 slots
 * each (the reference's in buffer has 3; see the header). Returns the vertex
 * count; *trivial is set when the trivial-accept path was taken. */
static int clip_triangle(ho_v2f* in, ho_v2f* out, int* trivial) {
    if (is_vertex_visible(in[0].clip_pos) && is_vertex_visible(in[1].clip_pos) &&
        is_vertex_visible(in[2].clip_pos)) {
        out[0] = in[0]; out[1] = in[1]; out[2] = in[2];
        *trivial = 1;
        return 3;
    }
    *trivial = 0;
    int n = 3;
    n = clip_against_plane(PW, n, in, out); if (n < 3) return 0;
    n = clip_against_plane(PX, n, out, in); if (n < 3) return 0;
    n = clip_against_plane(NX, n, in, out); if (n < 3) return 0;
    n = clip_against_plane(PY, n, out, in); if (n < 3) return 0;
    n = clip_against_plane(NY, n, in, out); if (n < 3) return 0;
    n = clip_against_plane(PZ, n, out, in); if (n < 3) return 0;
    n = clip_against_plane(NZ, n, in, out); if (n < 3) return 0;
    return n;
}
int ho_clip_triangle(const ho_v2f* in3, ho_v2f* out) {
    ho_v2f a[30], b[30];
    int trivial;
    a[0] = in3[0]; a[1] = in3[1]; a[2] = in3[2];
    int n = clip_triangle(a, b, &trivial);
    for (int i = 0; i < n; i++) out[i] = b[i];
    return n;
}

/* is_back_facing (SR/graphics.cpp:172-180). */
static int back_facing(const float* a, const float* b, const float* c) {
    float signed_area = a[0] * b[1] - a[1] * b[0] + b[0] * c[1] - b[1] * c[0] + c[0] * a[1] - c[1] * a[0];
    return signed_area <= 0;
}
int ho_is_back_facing(const float* ndc9) { return back_facing(ndc9, ndc9 + 3, ndc9 + 6); }

/* barycentric (SR/graphics.cpp:222-233). */
static void barycentric(const float* A, const float* B, const float* C, const float* P, float* w) {
    float s0[3], s1[3];
    s0[0] = C[0] - A[0]; s0[1] = B[0] - A[0]; s0[2] = A[0] - P[0];
    s1[0] = C[1] - A[1]; s1[1] = B[1] - A[1]; s1[2] = A[1] - P[1];
    float ux = s0[1] * s1[2] - s0[2] * s1[1];
    float uy = s0[2] * s1[0] - s0[0] * s1[2];
    float uz = s0[0] * s1[1] - s0[1] * s1[0];
    if ((double)fabsf(uz) > 1e-2) {
        w[0] = 1.f - (ux + uy) / uz;
        w[1] = uy / uz;
        w[2] = ux / uz;
        return;
    }
    w[0] = -1; w[1] = 1; w[2] = 1;
}
void ho_barycentric(const float* A, const float* B, const float* C, const float* P, float* out) {
    barycentric(A, B, C, P, out);
}

typedef struct {
    const ctx_t* c;
    int kind;
    ho_target* rb;
    ho_stats* st;
    int32_t* winner;
} draw_t;

/* rasterize_triangle (SR/graphics.cpp:314-376). */
static void rasterize_triangle(draw_t* d, const ho_v2f* v, int32_t prim) {
    float ndc[3][3];
    for (int i = 0; i < 3; i++)
        for (int k = 3; k--;) ndc[i][k] = v[i].clip_pos[k] / v[i].clip_pos[3];
    d->st->triangles++;
    if (back_facing(ndc[0], ndc[1], ndc[2])) { d->st->culled++; return; }
    ho_target* rb = d->rb;
    float sc[3][2], sd[3], rw[3];
    for (int i = 0; i < 3; i++) {
        float win[3];
        ho_viewport_transform(rb->width, rb->height, ndc[i], win);
        sc[i][0] = win[0]; sc[i][1] = win[1]; sd[i] = win[2];
    }
    for (int i = 0; i < 3; i++) rw[i] = 1 / v[i].clip_pos[3];
    float bbmin[2] = {FLT_MAX, FLT_MAX}, bbmax[2] = {-FLT_MAX, -FLT_MAX};
    float clampv[2] = {(float)(rb->width - 1), (float)(rb->height - 1)};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 2; j++) {
            bbmin[j] = smax(0.f, smin(bbmin[j], sc[i][j]));
            bbmax[j] = smin(clampv[j], smax(bbmax[j], sc[i][j]));
        }
    int x0 = f2i(bbmin[0]), y0 = f2i(bbmin[1]);
    /* bbmin >= 0 always; a negative start means bbmin did not fit an int
     * (undefined in the reference, which would scan from INT_MIN): no pixels. */
    if (x0 < 0 || y0 < 0) return;
    for (int px = x0; (float)px <= bbmax[0]; px++) {
        for (int py = y0; (float)py <= bbmax[1]; py++) {
            d->st->bbox_pixels++;
            float P[2] = {(float)px, (float)py}, w[3];
            barycentric(sc[0], sc[1], sc[2], P, w);
            if (w[0] < 0 || w[1] < 0 || w[2] < 0) continue;
            float z = 0.0f;   /* interpolate_depth: Vector3 dot, reverse order (SR/graphics.cpp:186-194) */
            z += sd[2] * w[2];
            z += sd[1] * w[1];
            z += sd[0] * w[0];
            size_t idx = (size_t)py * rb->width + px;
            if (z > rb->depth[idx]) continue;
            /* interpolate_varyings (SR/graphics.cpp:205-220) */
            ho_v2f iv;
            const float* s0 = (const float*)&v[0];
            const float* s1 = (const float*)&v[1];
            const float* s2 = (const float*)&v[2];
            float* dst = (float*)&iv;
            float w0 = rw[0] * w[0], w1 = rw[1] * w[1], w2 = rw[2] * w[2];
            float normalizer = 1 / (w0 + w1 + w2);
            for (int k = 0; k < 13; k++) {
                float sum = s0[k] * w0 + s1[k] * w1 + s2[k] * w2;
                dst[k] = sum * normalizer;
            }
            d->st->fragments++;
            rb->depth[idx] = z;                            /* RenderBuffer::set_depth */
            if (d->kind == HO_SHADER_SHADOW && d->c->u->shadow_fp32) {   /* extension: depth-only shadow pass */
                if (d->winner) d->winner[idx] = prim;
                continue;
            }
            rgba col = fragment_impl(d->c, d->kind, &iv);
            unsigned char* cp = rb->color + idx * 4;       /* RenderBuffer::set_color (SR/renderbuffer.cpp:38-44) */
            cp[0] = (unsigned char)(col.r * 255);
            cp[1] = (unsigned char)(col.g * 255);
            cp[2] = (unsigned char)(col.b * 255);
            if (d->winner) d->winner[idx] = prim;
        }
    }
}

/* graphics_draw_triangle (SR/graphics.cpp:378-407). */
int ho_draw(const ho_mesh* mesh, int kind, const ho_uniforms* u, const ho_image* diffuse,
            const ho_image* normal, const ho_target* shadow_map, ho_target* target,
            ho_stats* stats, int32_t* winner) {
    ctx_t c;
    ho_stats local;
    memset(&local, 0, sizeof local);
    ctx_init(&c, u, diffuse, normal, shadow_map);
    draw_t d = {&c, kind, target, &local, winner};
    ho_v2f in[30], out[30];
    for (int32_t f = 0; f < mesh->nfaces; f++) {
        for (int j = 0; j < 3; j++) {
            ho_a2v a;
            int ip = mesh->idx_pos[3 * f + j];
            int in_ = mesh->idx_nrm ? mesh->idx_nrm[3 * f + j] : ip;
            int iu = mesh->idx_uv ? mesh->idx_uv[3 * f + j] : ip;
            memcpy(a.obj_pos, mesh->pos + 3 * (size_t)ip, 3 * sizeof(float));
            /* Model::normal normalises the stored normal in place (SR/model.cpp:108-111) */
            float* np = mesh->normals + 3 * (size_t)in_;
            if (!(mesh->flags & HO_MESH_FETCHED_NORMALS)) {
                v3 n = normalize3(mk3(np[0], np[1], np[2]));
                np[0] = n.x; np[1] = n.y; np[2] = n.z;
            }
            memcpy(a.obj_normal, np, 3 * sizeof(float));
            memcpy(a.uv, mesh->uvs + 2 * (size_t)iu, 2 * sizeof(float));
            in[j] = vertex_impl(&c, kind, &a);
        }
        local.faces++;
        int trivial;
        int n = clip_triangle(in, out, &trivial);
        if (!trivial) local.clipped_faces++;
        for (int j = 0; j < n - 2; j++) {
            ho_v2f tri[3] = {out[0], out[j + 1], out[j + 2]};
            rasterize_triangle(&d, tri, f * 16 + (trivial ? 0 : 1 + j));
        }
    }
    if (stats) *stats = local;
    return 0;
}

void ho_target_clear(ho_target* t, const float* c, float depth) {   /* SR/renderbuffer.cpp:59-76 */
    size_t n = (size_t)t->width * t->height;
    unsigned char b[4];
    for (int k = 0; k < 4; k++) {
        float v = c[k] * 255;
        b[k] = (v >= 0.0f && v < 256.0f) ? (unsigned char)v : 255;   /* alpha 255*255 defined as 255 */
    }
    for (size_t i = 0; i < n; i++) {
        memcpy(t->color + 4 * i, b, 4);
        t->depth[i] = depth;
    }
}

/* DrawModel::draw (SR/scene.h:53-99). */
int ho_render_frame(const ho_mesh* mesh, int kind, int enable_shadow, const ho_uniforms* u,
                    const ho_image* diffuse, const ho_image* normal, ho_target* shadow_map,
                    ho_target* target, ho_stats* shadow_stats, ho_stats* main_stats) {
    ho_uniforms uu = *u;
    uu.enable_shadow = enable_shadow;
    if (enable_shadow) ho_draw(mesh, HO_SHADER_SHADOW, &uu, diffuse, normal, NULL, shadow_map, shadow_stats, NULL);
    ho_draw(mesh, kind, &uu, diffuse, normal, enable_shadow ? shadow_map : NULL, target, main_stats, NULL);
    if (enable_shadow) {
        const float black[4] = {0, 0, 0, 1};
        ho_target_clear(shadow_map, black, FLT_MAX);
    }
    return 0;
}
