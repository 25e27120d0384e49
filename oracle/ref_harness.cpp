// ref_harness.cpp -- builds the reference's own rasterisation path into
// oracle/_ref/libhanaref.so (TEST INFRASTRUCTURE ONLY; see oracle/Makefile).
//
// This file holds no reference code.  It #includes the reference's .cpp files
// from where they lie (the include path points into the reference tree) as one
// translation unit, so that the reference's static functions in graphics.cpp
// (clip_triangle, rasterize_triangle, barycentric, ...) are reachable, and adds
// a thin extern "C" driver with the same entry points and POD types as the
// oracle restatement (hana_oracle.h, prefix ref_ instead of ho_).
//
// Driver notes:
//  * graphics_draw_triangle (SR/graphics.cpp:378-407) passes its 3-slot v2fs
//    array as the input ping-pong buffer of clip_triangle, which writes up to 10
//    vertices into it for triangles that cross the +X/+Y/+Z planes (stack
//    overflow).  ref_draw therefore runs its own face loop: the same calls to
//    Model-style attribute fetches, IShader::vertex, clip_triangle and
//    rasterize_triangle, with 30 slots on both sides.  ref_draw_model calls the
//    unmodified graphics_draw_triangle on a Model read from an OBJ file, for
//    scenes without clipped triangles.
//  * ref_build_uniforms builds the camera matrices from the reference's
//    lookat/perspective with the constants of SR/camera.cpp:11-14, as
//    Camera::get_view_matrix/get_proj_matrix do.  camera.cpp and scene.cpp are
//    compiled too, for the Camera and Scene state entry points; camera.cpp's
//    input callbacks (get_cursor_pos, scroll_callback, button_callback,
//    update_camera) call the Win32 platform layer, are referenced by nothing
//    here, and are dropped at link time (oracle/ref.map exports only the ref_
//    entry points; --gc-sections).
//  * DrawModel::draw (SR/scene.h:53-99) is restated in ref_build_uniforms and
//    ref_render_frame with the reference's types and functions.
//  * Built with -DHANA_DROPIN (oracle/_ref/libhanaref_dropin.so), the
//    reference's graphics_draw_triangle is replaced by the MI355X drop-in, so
//    ref_draw_model runs the reference's Model, shaders and RenderBuffer
//    through libhana_gpu.so; every other entry point stays on the CPU.
#include <cfloat>
#include <cstring>
#include <limits>

#include "vector.cpp"   // first: its specialisations must precede first use
#ifdef HANA_DROPIN
// The drop-in build keeps the CPU face loop under another name and takes
// graphics_draw_triangle from hana-softwarerenderer_amd/dropin/graphics_hana.cpp.
#define graphics_draw_triangle graphics_draw_triangle_cpu
#include "graphics.cpp"
#undef graphics_draw_triangle
void graphics_draw_triangle(DrawData* draw_data);
#else
#include "graphics.cpp"
#endif
#include "IShader.cpp"
#include "renderbuffer.cpp"
#include "color.cpp"
#include "maths.cpp"
#include "tgaimage.cpp"
#include "model.cpp"
#include "gameobject.cpp"
#include "scene.h"      // AMBIENT, LightColor (SR/scene.h:8-9)
#include "camera.cpp"
#include "scene.cpp"
#ifdef HANA_DROPIN
#include "graphics_hana.cpp"
#endif

#include "hana_oracle.h"

namespace {

Matrix4x4 to_mat(const float* p) {
    Matrix4x4 m;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) m[i][j] = p[4 * i + j];
    return m;
}
void from_mat(const Matrix4x4& m, float* p) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) p[4 * i + j] = m[i][j];
}
Vector3f v3(const float* p) { return Vector3f(p[0], p[1], p[2]); }
Color col(const float* p) { return Color(p[0], p[1], p[2], p[3]); }

IShader* make_shader(int kind) {
    switch (kind) {
    case HO_SHADER_GROUND: return new GroundShader();
    case HO_SHADER_TOON: return new ToonShader();
    case HO_SHADER_TEXTURE: return new TextureShader();
    case HO_SHADER_TEXTURE_LIGHT: return new TextureWithLightShader();
    case HO_SHADER_BLINN: return new BlinnShader();
    case HO_SHADER_NORMALMAP: return new NormalMapShader();
    case HO_SHADER_SHADOW: return new ShadowShader();
    default: return nullptr;
    }
}

// Wraps a shader to count fragment() invocations.
struct Counting : IShader {
    IShader* inner;
    long long fragments = 0;
    explicit Counting(IShader* s) : inner(s) {}
    shader_struct_v2f vertex(shader_struct_a2v* a) override { return inner->vertex(a); }
    bool fragment(shader_struct_v2f* v, Color& c) override {
        fragments++;
        return inner->fragment(v, c);
    }
};

TGAImage* make_image(const ho_image* im) {
    if (!im || !im->data) return new TGAImage();
    TGAImage* t = new TGAImage(im->width, im->height, im->bytespp);
    memcpy(t->buffer(), im->data, (size_t)im->width * im->height * im->bytespp);
    return t;
}

RenderBuffer* make_rb(const ho_target* t) {
    RenderBuffer* rb = new RenderBuffer(t->width, t->height);
    size_t n = (size_t)t->width * t->height;
    memcpy(rb->color_buffer, t->color, n * 4);
    memcpy(rb->depth_buffer, t->depth, n * sizeof(float));
    return rb;
}
void store_rb(const RenderBuffer* rb, ho_target* t) {
    size_t n = (size_t)t->width * t->height;
    memcpy(t->color, rb->color_buffer, n * 4);
    memcpy(t->depth, rb->depth_buffer, n * sizeof(float));
}

struct Bound {
    Material mat{};
    ShaderData sd{};
    TGAImage* diffuse = nullptr;
    TGAImage* normal = nullptr;
    RenderBuffer* shadow = nullptr;
    Bound(const ho_uniforms* u, const ho_image* d, const ho_image* n, const ho_target* sm) {
        diffuse = make_image(d);
        normal = make_image(n);
        mat.diffuse_map = diffuse;
        mat.normal_map = normal;
        mat.specular_map = nullptr;
        mat.color = col(u->mat_color);
        mat.specular = col(u->mat_specular);
        mat.gloss = u->gloss;
        mat.bump_scale = u->bump_scale;
        sd.matrial = &mat;
        sd.targetBuffer = nullptr;
        shadow = sm ? make_rb(sm) : nullptr;
        sd.shadow_map = shadow;
        sd.enable_shadow = u->enable_shadow != 0;
        sd.view_Pos = v3(u->view_pos);
        sd.light_dir = v3(u->light_dir);
        sd.light_color = col(u->light_color);
        sd.ambient = col(u->ambient);
        sd.model_matrix = to_mat(u->model);
        sd.model_matrix_I = to_mat(u->model_inv);
        sd.view_matrix = to_mat(u->view);
        sd.projection_matrix = to_mat(u->proj);
        sd.light_vp_matrix = to_mat(u->light_vp);
        sd.camera_vp_matrix = to_mat(u->camera_vp);
    }
    ~Bound() {
        delete diffuse;
        delete normal;
        delete shadow;
    }
};

shader_struct_v2f to_v2f(const ho_v2f* v) {
    shader_struct_v2f o;
    static_assert(sizeof(shader_struct_v2f) == 13 * sizeof(float), "v2f layout");
    memcpy(&o, v, sizeof o);
    return o;
}

}  // namespace

extern "C" {

void ref_mat4_mul(const float* a, const float* b, float* out) { from_mat(to_mat(a) * to_mat(b), out); }
void ref_mat4_mul_vec4(const float* m, const float* v, float* out) {
    Vector4f r = to_mat(m) * Vector4f(v[0], v[1], v[2], v[3]);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
void ref_mat4_invert(const float* m, float* out) { from_mat(to_mat(m).invert(), out); }
void ref_translate(float tx, float ty, float tz, float* out) { from_mat(translate(tx, ty, tz), out); }
void ref_scale(float sx, float sy, float sz, float* out) { from_mat(scale(sx, sy, sz), out); }
void ref_rotate_x(float a, float* out) { from_mat(rotate_x(a), out); }
void ref_rotate_y(float a, float* out) { from_mat(rotate_y(a), out); }
void ref_rotate_z(float a, float* out) { from_mat(rotate_z(a), out); }
void ref_rotate(float a, float vx, float vy, float vz, float* out) { from_mat(rotate(a, vx, vy, vz), out); }
void ref_lookat(const float* e, const float* t, const float* u, float* out) { from_mat(lookat(v3(e), v3(t), v3(u)), out); }
void ref_orthographic(float r, float t, float n, float f, float* out) { from_mat(orthographic(r, t, n, f), out); }
void ref_perspective(float fovy, float aspect, float n, float f, float* out) {
    from_mat(perspective(fovy, aspect, n, f), out);
}
void ref_viewport_transform(int w, int h, const float* ndc, float* out) {
    Vector3f r = viewport_transform(w, h, v3(ndc));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void ref_model_matrix(const float* p, const float* r, const float* s, float* out) {
    GameObject go(v3(p), v3(r), v3(s));
    from_mat(go.GetModelMatrix(), out);
}

// DrawModel::draw lines 55-71 (SR/scene.h) with Camera(pos, target, aspect)'s
// get_view_matrix / get_proj_matrix (SR/camera.cpp:82-92) and the scene's
// material (SR/scene.cpp:81-84).
void ref_build_uniforms(const float* cam_pos, const float* cam_target, float cam_aspect,
                        const float* light_pos, const float* obj_position,
                        const float* obj_rotation_deg, const float* obj_scale, int fb_w, int fb_h,
                        int enable_shadow, ho_uniforms* u) {
    static const float NEAR_ = 0.1f;
    static const float FAR_ = 10000;
    static const float FOVY_ = TO_RADIANS(60);
    static const Vector3f UP_ = {0, 1, 0};
    memset(u, 0, sizeof *u);
    Vector3f position = v3(cam_pos), target = v3(cam_target);
    Vector3f up = Vector3f(UP_.x, UP_.y, UP_.z);
    Matrix4x4 view_matrix = lookat(position, target, up);
    Matrix4x4 projection_matrix = perspective(FOVY_, cam_aspect, NEAR_, FAR_);
    GameObject go(v3(obj_position), v3(obj_rotation_deg), v3(obj_scale));
    Matrix4x4 model_matrix = go.GetModelMatrix();
    Matrix4x4 model_matrix_I = model_matrix.invert();
    GameObject light(v3(light_pos));
    Vector3f target_position = Vector3f::Zero;   // Camera::get_target_position
    Vector3f light_dir = (light.transform.position - target_position).normalize();
    float aspect = (float)fb_w / (float)fb_h;
    Matrix4x4 light_vp = orthographic(aspect, 1, 0, 5) * lookat(light.transform.position, target_position, Vector3f(0, 1, 0));
    Matrix4x4 camera_vp = projection_matrix * view_matrix;
    from_mat(model_matrix, u->model);
    from_mat(model_matrix_I, u->model_inv);
    from_mat(view_matrix, u->view);
    from_mat(projection_matrix, u->proj);
    from_mat(light_vp, u->light_vp);
    from_mat(camera_vp, u->camera_vp);
    u->view_pos[0] = position.x; u->view_pos[1] = position.y; u->view_pos[2] = position.z;
    u->light_dir[0] = light_dir.x; u->light_dir[1] = light_dir.y; u->light_dir[2] = light_dir.z;
    Color lc = LightColor, am = AMBIENT, wc = Color::White;
    float* dst[4] = {u->light_color, u->ambient, u->mat_color, u->mat_specular};
    Color src[4] = {lc, am, wc, wc};
    for (int k = 0; k < 4; k++) {
        dst[k][0] = src[k].r; dst[k][1] = src[k].g; dst[k][2] = src[k].b; dst[k][3] = src[k].a;
    }
    u->gloss = 50;
    u->bump_scale = 1;
    u->enable_shadow = enable_shadow;
}

void ref_barycentric(const float* A, const float* B, const float* C, const float* P, float* out) {
    Vector3f w = barycentric(Vector2f(A[0], A[1]), Vector2f(B[0], B[1]), Vector2f(C[0], C[1]), Vector2f(P[0], P[1]));
    out[0] = w.x; out[1] = w.y; out[2] = w.z;
}

int ref_clip_triangle(const ho_v2f* in3, ho_v2f* out) {
    shader_struct_v2f in[30], res[30];
    for (int i = 0; i < 3; i++) in[i] = to_v2f(&in3[i]);
    int n = clip_triangle(0, in, res);
    for (int i = 0; i < n; i++) memcpy(&out[i], &res[i], sizeof(ho_v2f));
    return n;
}

int ref_is_back_facing(const float* ndc9) {
    Vector3f n[3] = {v3(ndc9), v3(ndc9 + 3), v3(ndc9 + 6)};
    return is_back_facing(n) ? 1 : 0;
}

ho_v2f ref_vertex(int kind, const ho_uniforms* u, const ho_a2v* a) {
    Bound b(u, nullptr, nullptr, nullptr);
    IShader* s = make_shader(kind);
    s->shader_data = &b.sd;
    shader_struct_a2v in;
    in.obj_pos = v3(a->obj_pos);
    in.obj_normal = v3(a->obj_normal);
    in.uv = Vector2f(a->uv[0], a->uv[1]);
    shader_struct_v2f o = s->vertex(&in);
    if (kind != HO_SHADER_GROUND && kind != HO_SHADER_TOON) o.intensity = 0;   // left uninitialised by the shader
    ho_v2f r;
    memcpy(&r, &o, sizeof r);
    delete s;
    return r;
}

int ref_fragment(int kind, const ho_uniforms* u, const ho_image* diffuse, const ho_image* normal,
                 const ho_target* shadow_map, const ho_v2f* v2f, float* out) {
    Bound b(u, diffuse, normal, shadow_map);
    IShader* s = make_shader(kind);
    s->shader_data = &b.sd;
    shader_struct_v2f v = to_v2f(v2f);
    Color c;
    s->fragment(&v, c);
    out[0] = c.r; out[1] = c.g; out[2] = c.b; out[3] = c.a;
    delete s;
    return 0;
}

// Face loop of graphics_draw_triangle with 30-slot clip buffers on both sides.
int ref_draw(const ho_mesh* mesh, int kind, const ho_uniforms* u, const ho_image* diffuse,
             const ho_image* normal, const ho_target* shadow_map, ho_target* target, ho_stats* stats,
             int32_t* winner) {
    (void)winner;
    Bound b(u, diffuse, normal, shadow_map);
    IShader* inner = make_shader(kind);
    inner->shader_data = &b.sd;
    Counting shader(inner);
    shader.shader_data = &b.sd;
    RenderBuffer* rb = make_rb(target);
    DrawData dd;
    dd.model = nullptr;
    dd.shader = &shader;
    dd.render_buffer = rb;
    long long clipped = 0;
    shader_struct_v2f v2fs[30];
    for (int i = 0; i < mesh->nfaces; i++) {
        for (int j = 0; j < 3; j++) {
            int ip = mesh->idx_pos[3 * i + j];
            int in_ = mesh->idx_nrm ? mesh->idx_nrm[3 * i + j] : ip;
            int iu = mesh->idx_uv ? mesh->idx_uv[3 * i + j] : ip;
            shader_struct_a2v a2v;
            a2v.obj_pos = v3(mesh->pos + 3 * (size_t)ip);
            Vector3f n = v3(mesh->normals + 3 * (size_t)in_);   // Model::normal: normalise in place
            if (mesh->flags & HO_MESH_FETCHED_NORMALS) {
                a2v.obj_normal = n;
            } else {
                a2v.obj_normal = n.normalize();
                mesh->normals[3 * (size_t)in_ + 0] = n.x;
                mesh->normals[3 * (size_t)in_ + 1] = n.y;
                mesh->normals[3 * (size_t)in_ + 2] = n.z;
            }
            a2v.uv = Vector2f(mesh->uvs[2 * (size_t)iu], mesh->uvs[2 * (size_t)iu + 1]);
            v2fs[j] = dd.shader->vertex(&a2v);
        }
        if (!(is_vertex_visible(v2fs[0].clip_pos) && is_vertex_visible(v2fs[1].clip_pos) &&
              is_vertex_visible(v2fs[2].clip_pos)))
            clipped++;
        shader_struct_v2f clip_v2fs[3 * 10];
        int num_vertices = clip_triangle(0, v2fs, clip_v2fs);
        for (int j = 0; j < num_vertices - 2; j++) {
            shader_struct_v2f ret_v2fs[3];
            ret_v2fs[0] = clip_v2fs[0];
            ret_v2fs[1] = clip_v2fs[j + 1];
            ret_v2fs[2] = clip_v2fs[j + 2];
            rasterize_triangle(&dd, ret_v2fs);
        }
    }
    store_rb(rb, target);
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->faces = mesh->nfaces;
        stats->clipped_faces = clipped;
        stats->fragments = shader.fragments;
    }
    delete rb;
    delete inner;
    return 0;
}

// The unmodified entry point on a Model read from an OBJ file by the
// reference's own parser (SR/model.cpp:6-48).
int ref_draw_model(const char* obj_path, int kind, const ho_uniforms* u, const ho_image* diffuse,
                   const ho_image* normal, const ho_target* shadow_map, ho_target* target) {
    Bound b(u, diffuse, normal, shadow_map);
    Model model(obj_path);
    IShader* shader = make_shader(kind);
    shader->shader_data = &b.sd;
    RenderBuffer* rb = make_rb(target);
    DrawData dd;
    dd.model = &model;
    dd.shader = shader;
    dd.render_buffer = rb;
    graphics_draw_triangle(&dd);
    store_rb(rb, target);
    delete rb;
    delete shader;
    return model.nfaces();
}

// What graphics_draw_triangle's face loop fetches from a Model read from an
// OBJ file (SR/graphics.cpp:380-386): per corner, in order, vert / normal / uv.
// Returns the face count; fills at most max_faces faces.
int ref_model_corners(const char* obj_path, float* pos, float* nrm, float* uv, int max_faces) {
    Model model(obj_path);
    int nf = model.nfaces();
    for (int i = 0; i < nf && i < max_faces; i++)
        for (int j = 0; j < 3; j++) {
            size_t c = 3 * (size_t)i + j;
            Vector3f p = model.vert(i, j);
            Vector3f n = model.normal(i, j);
            Vector2f t = model.uv(i, j);
            for (int k = 0; k < 3; k++) {
                pos[3 * c + k] = p[k];
                nrm[3 * c + k] = n[k];
            }
            uv[2 * c] = t[0];
            uv[2 * c + 1] = t[1];
        }
    return nf;
}

// TGAImage::read_tga_file (+ Model::load_texture's flip when flip_v).  Returns
// 1 on success and copies at most cap bytes.
int ref_tga_read(const char* path, int flip_v, int* w, int* h, int* bpp, unsigned char* out, size_t cap) {
    TGAImage img;
    if (!img.read_tga_file(path)) return 0;
    if (flip_v) img.flip_vertically();
    *w = img.get_width();
    *h = img.get_height();
    *bpp = img.get_bytespp();
    size_t n = (size_t)*w * *h * *bpp;
    memcpy(out, img.buffer(), n < cap ? n : cap);
    return 1;
}

// TGAImage::write_tga_file of an image with the given bytes.
int ref_tga_write(const char* path, int w, int h, int bpp, const unsigned char* bytes, int rle) {
    TGAImage img(w, h, bpp);
    memcpy(img.buffer(), bytes, (size_t)w * h * bpp);
    return img.write_tga_file(path, rle != 0) ? 1 : 0;
}

void ref_target_clear(ho_target* t, const float* c, float depth) {
    RenderBuffer* rb = make_rb(t);
    rb->renderbuffer_clear_color(Color(c[0], c[1], c[2], c[3]));
    rb->renderbuffer_clear_depth(depth);
    store_rb(rb, t);
    delete rb;
}

// DrawModel::draw (SR/scene.h:73-98): shadow pass, main pass, shadow clear.
int ref_render_frame(const ho_mesh* mesh, int kind, int enable_shadow, const ho_uniforms* u,
                     const ho_image* diffuse, const ho_image* normal, ho_target* shadow_map,
                     ho_target* target, ho_stats* shadow_stats, ho_stats* main_stats) {
    ho_uniforms uu = *u;
    uu.enable_shadow = enable_shadow;
    if (enable_shadow) ref_draw(mesh, HO_SHADER_SHADOW, &uu, diffuse, normal, nullptr, shadow_map, shadow_stats, nullptr);
    ref_draw(mesh, kind, &uu, diffuse, normal, enable_shadow ? shadow_map : nullptr, target, main_stats, nullptr);
    if (enable_shadow) {
        // The reference clears with Color::Black, whose alpha (255) times 255
        // does not fit a byte; the clear here uses alpha 1 (-> 255) instead,
        // RGB and depth are the reference's (SR/scene.h:96-97).
        const float black[4] = {0, 0, 0, 1};
        ref_target_clear(shadow_map, black, std::numeric_limits<float>::max());
    }
    return 0;
}

// Camera::update_transform (SR/camera.cpp:63-70) on Camera(pos, target, aspect)
// for Motion {orbit, pan, dolly} = motion[0..1], [2..3], [4].  Camera keeps
// its target private with no accessor (get_target_position returns the
// origin), so position and target are read through the class's layout.
void ref_camera_update(const float* pos, const float* target, float aspect, const float* motion, float* out_pos,
                       float* out_target) {
    Camera cam(v3(pos), v3(target), aspect);
    Motion m;
    m.orbit = Vector2f(motion[0], motion[1]);
    m.pan = Vector2f(motion[2], motion[3]);
    m.dolly = motion[4];
    cam.update_transform(m);
    struct Layout {
        Vector3f position, target;
        float aspect;
    };
    static_assert(sizeof(Layout) == sizeof(Camera), "Camera holds position, target, aspect");
    Layout l;
    memcpy(&l, &cam, sizeof l);
    out_pos[0] = l.position.x; out_pos[1] = l.position.y; out_pos[2] = l.position.z;
    out_target[0] = l.target.x; out_target[1] = l.target.y; out_target[2] = l.target.z;
}

// SingleModelScene(obj_path) over a w x h RenderBuffer, driven by n pressed
// keys (SR/scene.cpp:45-63,127-147).  After each key the scene ticks (it draws
// its frame) and out[5*i..] receives the light position, enable_shadow and the
// shader index, read from the "shader: " line tick() writes (SR/scene.cpp:119).
// Returns 0, or -1 when a shader line is not recognised.
int ref_scene_keys(const char* obj_path, int w, int h, const int* keys, int n, float* out) {
    static const char* des[6] = {"6.NormalMap", "5.Blinn", "4.Texture_Light(no shadow)", "5.Texture(no shadow)",
                                 "2.Toon(no shadow)", "1.Ground(no shadow)"};   // SR/scene.cpp:86-97
    RenderBuffer rb(w, h);
    SingleModelScene scene(obj_path, &rb);
    for (int i = 0; i < n; i++) {
        scene.on_key_input((keycode_t)keys[i], 1);
        scene.tick(0.0f);
        const char* line = strstr(scene.get_text(), "shader: ");
        int idx = -1;
        for (int k = 0; line && k < 6 && idx < 0; k++) {
            size_t len = strlen(des[k]);
            if (strncmp(line + 8, des[k], len) == 0 && line[8 + len] == '\n') idx = k;
        }
        if (idx < 0) return -1;
        Vector3f lp = scene.light->transform.position;
        float* o = out + 5 * i;
        o[0] = lp.x; o[1] = lp.y; o[2] = lp.z;
        o[3] = scene.enable_shadow ? 1.0f : 0.0f;
        o[4] = (float)idx;
    }
    return 0;
}

}  // extern "C"
