// graphics_hana.cpp -- graphics_draw_triangle (SR/graphics.cpp:378-407) on the
// MI355X through libhana_gpu.so (include/hana_gpu.h).
//
// A C++ project built on the renderer's own headers compiles this file in place
// of the definition of graphics_draw_triangle in graphics.cpp and links
// libhana_gpu.so.  Callers do not change: the same DrawData
// (SR/graphics.h:9-13), the shader subclasses of SR/IShader.h:132-165 bound to
// their ShaderData, a Model, and a host RenderBuffer whose colour (RGB) and
// depth are updated in place as RenderBuffer::set_color / set_depth would
// (SR/renderbuffer.cpp:26-45; alpha is left as it was).
//
// Each call fetches every corner through Model::vert / normal / uv in the
// face loop's order (Model::normal renormalises its stored vector at every
// fetch, SR/model.cpp:108-111, so the order of fetches is part of the result),
// uploads the corners, the material textures, the target and a sampled shadow
// map, runs one pass, and copies colour and depth back.  Nothing is cached
// between calls, so a Model, TGAImage or RenderBuffer may change or be freed
// between calls as with the CPU function.  Scenes that stay resident belong on
// hana_render_frame / hana_batch_* instead.
//
// Errors: the CPU function has no error path.  A failure here (no device, a
// shader subclass without a device program, a device error) prints its cause
// to stderr and aborts; there is no CPU fallback.  A clipped polygon of more
// than 16 vertices (possible only through rounding) is dropped with a warning.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "graphics.h"
#include "hana_gpu.h"

namespace hana_dropin {

struct Device {
    hana_ctx* ctx = nullptr;
    hana_target* target = nullptr;
    int tw = 0, th = 0;
    hana_target* shadow = nullptr;
    int sw = 0, sh = 0;
    std::vector<float> pos, nrm, uv, depth;
    std::vector<int32_t> idx;
    std::vector<uint8_t> rgba;
};

[[noreturn]] void die(hana_ctx* ctx, const char* what) {
    std::fprintf(stderr, "graphics_draw_triangle: %s (%s)\n", what, ctx ? hana_gpu_last_error(ctx) : "no device");
    std::fflush(stderr);
    std::abort();
}

void check(hana_ctx* ctx, int rc, const char* what) {
    if (rc != HANA_OK) die(ctx, what);
}

Device& device() {
    static Device d;
    if (!d.ctx) check(nullptr, hana_gpu_create(0, &d.ctx), "no HIP device for libhana_gpu.so");
    return d;
}

// The shader subclass decides the device program; a subclass outside
// SR/IShader.h:132-165 has none.
int shader_kind(IShader* s) {
    if (dynamic_cast<GroundShader*>(s)) return HANA_SHADER_GROUND;
    if (dynamic_cast<ToonShader*>(s)) return HANA_SHADER_TOON;
    if (dynamic_cast<TextureShader*>(s)) return HANA_SHADER_TEXTURE;
    if (dynamic_cast<TextureWithLightShader*>(s)) return HANA_SHADER_TEXTURE_LIGHT;
    if (dynamic_cast<BlinnShader*>(s)) return HANA_SHADER_BLINN;
    if (dynamic_cast<NormalMapShader*>(s)) return HANA_SHADER_NORMALMAP;
    if (dynamic_cast<ShadowShader*>(s)) return HANA_SHADER_SHADOW;
    return -1;
}

void put_mat(const Matrix4x4& m, float* out) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = m[i][j];
}

void put_color(const Color& c, float* out) {
    out[0] = c.r;
    out[1] = c.g;
    out[2] = c.b;
    out[3] = c.a;
}

// ShaderData + Material (SR/IShader.h:7-32) as the device reads them.
hana_uniforms uniforms_of(const ShaderData* sd) {
    hana_uniforms u;
    std::memset(&u, 0, sizeof u);
    put_mat(sd->model_matrix, u.model);
    put_mat(sd->model_matrix_I, u.model_inv);
    put_mat(sd->view_matrix, u.view);
    put_mat(sd->projection_matrix, u.proj);
    put_mat(sd->light_vp_matrix, u.light_vp);
    put_mat(sd->camera_vp_matrix, u.camera_vp);
    for (int k = 0; k < 3; k++) {
        u.view_pos[k] = sd->view_Pos[k];
        u.light_dir[k] = sd->light_dir[k];
    }
    put_color(sd->light_color, u.light_color);
    put_color(sd->ambient, u.ambient);
    if (sd->matrial) {
        put_color(sd->matrial->color, u.mat_color);
        put_color(sd->matrial->specular, u.mat_specular);
        u.gloss = sd->matrial->gloss;
        u.bump_scale = sd->matrial->bump_scale;
    }
    // IShader::is_in_shadow samples only when both are set (SR/IShader.h:108)
    u.enable_shadow = (sd->enable_shadow && sd->shadow_map) ? 1 : 0;
    return u;
}

hana_texture* upload_texture(hana_ctx* ctx, TGAImage* t) {
    if (!t || !t->buffer() || t->get_width() <= 0 || t->get_height() <= 0) return nullptr;   // get() -> TGAColor()
    hana_texture* out = nullptr;
    check(ctx, hana_texture_create(ctx, t->get_width(), t->get_height(), t->get_bytespp(), t->buffer(), &out),
          "texture upload failed");
    return out;
}

hana_target* sized(hana_ctx* ctx, hana_target*& t, int& w, int& h, int want_w, int want_h) {
    if (t && (w != want_w || h != want_h)) {
        hana_target_destroy(ctx, t);
        t = nullptr;
    }
    if (!t) {
        check(ctx, hana_target_create(ctx, want_w, want_h, &t), "render target allocation failed");
        w = want_w;
        h = want_h;
    }
    return t;
}

void draw(DrawData* dd) {
    Device& d = device();
    hana_ctx* ctx = d.ctx;
    if (!dd || !dd->model || !dd->shader || !dd->shader->shader_data || !dd->render_buffer)
        die(ctx, "incomplete DrawData");
    int kind = shader_kind(dd->shader);
    if (kind < 0) die(ctx, "this IShader subclass has no device program");
    ShaderData* sd = dd->shader->shader_data;
    RenderBuffer* rb = dd->render_buffer;
    hana_uniforms u = uniforms_of(sd);
    if (u.enable_shadow && sd->shadow_map == rb) die(ctx, "the shadow map is the render target");

    // the face loop's vertex fetches (SR/graphics.cpp:380-386), one vertex per corner
    Model* m = dd->model;
    int F = m->nfaces();
    d.pos.resize(9 * (size_t)F);
    d.nrm.resize(9 * (size_t)F);
    d.uv.resize(6 * (size_t)F);
    d.idx.resize(3 * (size_t)F);
    for (int i = 0; i < F; i++)
        for (int j = 0; j < 3; j++) {
            size_t c = 3 * (size_t)i + j;
            Vector3f p = m->vert(i, j);
            Vector3f n = m->normal(i, j);
            Vector2f t = m->uv(i, j);
            for (int k = 0; k < 3; k++) {
                d.pos[3 * c + k] = p[k];
                d.nrm[3 * c + k] = n[k];
            }
            d.uv[2 * c] = t[0];
            d.uv[2 * c + 1] = t[1];
            d.idx[c] = (int32_t)c;
        }
    hana_mesh* mesh = nullptr;
    check(ctx, hana_mesh_create(ctx, 3 * F, d.pos.data(), d.nrm.data(), d.uv.data(), F, d.idx.data(), &mesh),
          "mesh upload failed");
    hana_texture* diffuse = sd->matrial ? upload_texture(ctx, sd->matrial->diffuse_map) : nullptr;
    hana_texture* normal = sd->matrial ? upload_texture(ctx, sd->matrial->normal_map) : nullptr;

    hana_target* tgt = sized(ctx, d.target, d.tw, d.th, rb->width, rb->height);
    hana_target* shadow = nullptr;
    if (u.enable_shadow) {
        RenderBuffer* s = sd->shadow_map;
        shadow = sized(ctx, d.shadow, d.sw, d.sh, s->width, s->height);
        check(ctx, hana_target_upload(ctx, shadow, s->color_buffer, s->depth_buffer), "shadow map upload failed");
    }
    size_t n = (size_t)rb->width * rb->height;
    d.rgba.resize(4 * n);
    d.depth.resize(n);
    hana_draw_desc desc;
    std::memset(&desc, 0, sizeof desc);
    desc.mesh = mesh;
    desc.shader = kind;
    desc.uniforms = &u;
    desc.diffuse = diffuse;
    desc.normal = normal;
    desc.shadow_map = shadow;
    desc.target = tgt;
    // a pass that outgrows the workspace is grown and rendered again inside
    // the read; a clipped polygon past the device's 16 vertices is dropped
    // with a warning (the reference's buffer holds 30)
    check(ctx, hana_target_upload(ctx, tgt, rb->color_buffer, rb->depth_buffer), "target upload failed");
    check(ctx, hana_draw(ctx, &desc), "draw failed");
    int rc = hana_target_read(ctx, tgt, d.rgba.data(), d.depth.data());
    if (rc == HANA_E_UNSUPPORTED)
        std::fprintf(stderr, "graphics_draw_triangle: %s\n", hana_gpu_last_error(ctx));
    else
        check(ctx, rc, "draw failed");
    for (size_t i = 0; i < n; i++) std::memcpy(rb->color_buffer + 4 * i, d.rgba.data() + 4 * i, 3);
    std::memcpy(rb->depth_buffer, d.depth.data(), sizeof(float) * n);
    if (diffuse) hana_texture_destroy(ctx, diffuse);
    if (normal) hana_texture_destroy(ctx, normal);
    hana_mesh_destroy(ctx, mesh);
}

}  // namespace hana_dropin

void graphics_draw_triangle(DrawData* draw_data) { hana_dropin::draw(draw_data); }
