/*
 * hana_oracle.h -- CPU restatement of the reference rasterisation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the GPU product is
 * compared against; it is never linked into, called by, or used as a fallback
 * of the product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * It restates, in plain C with the same floating-point operation order,
 * the functions of Hana-SoftwareRenderer listed below (SR/x = the reference's
 * Hana-SoftwareRenderer/x):
 *   graphics_draw_triangle / clip_triangle / rasterize_triangle  SR/graphics.cpp:7-407
 *   IShader helpers and the 7 shader programs                    SR/IShader.h:49-130, SR/IShader.cpp:5-180
 *   RenderBuffer colour/depth layout                              SR/renderbuffer.cpp:3-76
 *   TGAImage::get / TGAColor(p,bpp) / Color(TGAColor)             SR/tgaimage.cpp:248-253, SR/tgaimage.h:46-53, SR/color.cpp:5
 *   Vector / Matrix / Color arithmetic                            SR/vector.h, SR/matrix.h, SR/color.cpp
 *   transform builders, viewport                                  SR/maths.cpp:3-255
 *   camera matrices, model matrix, DrawModel::draw uniforms       SR/camera.cpp:11-14,82-92, SR/gameobject.cpp:12-17, SR/scene.h:53-99
 *
 * Build with -ffp-contract=off (no FMA contraction); every expression keeps
 * the reference's association order.
 */
#ifndef HANA_ORACLE_H
#define HANA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Shader kinds; numbering shared with the product's C ABI (hana_gpu.h). */
enum {
    HO_SHADER_GROUND = 0,        /* SR/IShader.cpp:5-15   */
    HO_SHADER_TOON = 1,          /* SR/IShader.cpp:23-39  */
    HO_SHADER_TEXTURE = 2,       /* SR/IShader.cpp:47-57  */
    HO_SHADER_TEXTURE_LIGHT = 3, /* SR/IShader.cpp:65-77  */
    HO_SHADER_BLINN = 4,         /* SR/IShader.cpp:85-109 */
    HO_SHADER_NORMALMAP = 5,     /* SR/IShader.cpp:117-162 */
    HO_SHADER_SHADOW = 6         /* SR/IShader.cpp:170-180 */
};

/* shader_struct_a2v (SR/IShader.h:35-39): 8 floats. */
typedef struct {
    float obj_pos[3];
    float obj_normal[3];
    float uv[2];
} ho_a2v;

/* shader_struct_v2f (SR/IShader.h:41-47): 13 floats, interpolated as a flat
 * float array by clipping and rasterisation (SR/graphics.cpp:7-17,205-220). */
typedef struct {
    float clip_pos[4];
    float world_pos[3];
    float world_normal[3];
    float uv[2];
    float intensity;
} ho_v2f;

/* Uniform block: ShaderData + Material scalars without pointers
 * (SR/IShader.h:7-32).  Matrices are row-major 4x4 (Matrix rows are
 * Vector<4>, SR/matrix.h:31).  Layout identical to hana_uniforms. */
typedef struct {
    float model[16];
    float model_inv[16];
    float view[16];
    float proj[16];
    float light_vp[16];
    float camera_vp[16];
    float view_pos[3];
    float light_dir[3];
    float light_color[4];
    float ambient[4];
    float mat_color[4];
    float mat_specular[4];
    float gloss;
    float bump_scale;
    int32_t enable_shadow;
    int32_t shadow_fp32;   /* fp32 shadow-map extension (no reference counterpart): see is_in_shadow */
} ho_uniforms;

/* TGAImage storage (SR/tgaimage.h:66-98): bytespp in {1,3,4}, bytes in
 * B,G,R,A order, texel (x,y) at (x + y*width)*bytespp. */
typedef struct {
    int32_t width, height, bytespp, pad_;
    const unsigned char* data;
} ho_image;

/* RenderBuffer (SR/renderbuffer.h:5-21): RGBA8 colour at (y*W+x)*4, fp32
 * depth at y*W+x. */
typedef struct {
    int32_t width, height;
    unsigned char* color;
    float* depth;
} ho_target;

/* Model accessors (SR/model.cpp:66-111) over index triplets.  normals are
 * normalised in place on every fetch, as Model::normal(iface, n) does, unless
 * flags has HO_MESH_FETCHED_NORMALS: then they already are the values the
 * fetches return (as an OBJ loader that replays the fetch sequence gives
 * them) and are used as stored. */
#define HO_MESH_FETCHED_NORMALS 1
typedef struct {
    int32_t nfaces, flags;
    const float* pos;          /* 3 per vertex */
    float* normals;            /* 3 per normal, mutated */
    const float* uvs;          /* 2 per uv */
    const int32_t* idx_pos;    /* 3 per face */
    const int32_t* idx_nrm;    /* 3 per face; NULL: use idx_pos */
    const int32_t* idx_uv;     /* 3 per face; NULL: use idx_pos */
} ho_mesh;

typedef struct {
    int64_t faces;          /* faces submitted */
    int64_t clipped_faces;  /* faces that failed the trivial-accept test */
    int64_t triangles;      /* fan triangles handed to rasterize_triangle */
    int64_t culled;         /* triangles rejected by is_back_facing */
    int64_t bbox_pixels;    /* pixels visited by the bbox scan */
    int64_t fragments;      /* fragment() invocations (the shaded-fragment count) */
} ho_stats;

/* ---- per-function entry points (known-answer tests) ---- */
void ho_mat4_mul(const float* a, const float* b, float* out);
void ho_mat4_mul_vec4(const float* m, const float* v, float* out);
void ho_mat4_invert(const float* m, float* out);
void ho_translate(float tx, float ty, float tz, float* out);
void ho_scale(float sx, float sy, float sz, float* out);
void ho_rotate_x(float angle, float* out);
void ho_rotate_y(float angle, float* out);
void ho_rotate_z(float angle, float* out);
void ho_rotate(float angle, float vx, float vy, float vz, float* out);
void ho_lookat(const float* eye, const float* target, const float* up, float* out);
void ho_orthographic(float right, float top, float n, float f, float* out);
void ho_perspective(float fovy, float aspect, float n, float f, float* out);
void ho_viewport_transform(int width, int height, const float* ndc, float* out);
void ho_model_matrix(const float* position, const float* rotation_deg, const float* scale3, float* out);
/* ShaderData as DrawModel::draw fills it (SR/scene.h:55-71) for a camera at
 * cam_pos looking at cam_target, a light at light_pos, a model transform and
 * a framebuffer of fb_w x fb_h. */
void ho_build_uniforms(const float* cam_pos, const float* cam_target, float cam_aspect,
                       const float* light_pos, const float* obj_position,
                       const float* obj_rotation_deg, const float* obj_scale,
                       int fb_w, int fb_h, int enable_shadow, ho_uniforms* out);

void ho_barycentric(const float* A, const float* B, const float* C, const float* P, float* out);
int ho_clip_triangle(const ho_v2f* in3, ho_v2f* out /* >= 10 */);
int ho_is_back_facing(const float* ndc9);
ho_v2f ho_vertex(int kind, const ho_uniforms* u, const ho_a2v* a2v);
/* Runs fragment() on a copy of *v2f; writes RGBA floats to out_rgba. */
int ho_fragment(int kind, const ho_uniforms* u, const ho_image* diffuse, const ho_image* normal,
                const ho_target* shadow_map, const ho_v2f* v2f, float* out_rgba);

/* ---- draw entry points ---- */
/* graphics_draw_triangle (SR/graphics.cpp:378-407) with a 30-slot clip
 * buffer on both sides.  winner (nullable, W*H) receives for every written
 * pixel the primitive id face*16 + (trivially accepted ? 0 : 1 + fan index). */
int ho_draw(const ho_mesh* mesh, int kind, const ho_uniforms* u, const ho_image* diffuse,
            const ho_image* normal, const ho_target* shadow_map, ho_target* target,
            ho_stats* stats, int32_t* winner);
/* DrawModel::draw (SR/scene.h:53-99): optional shadow pass into shadow_map,
 * main pass into target, then the shadow map clear (Black, FLT_MAX). */
int ho_render_frame(const ho_mesh* mesh, int kind, int enable_shadow, const ho_uniforms* u,
                    const ho_image* diffuse, const ho_image* normal, ho_target* shadow_map,
                    ho_target* target, ho_stats* shadow_stats, ho_stats* main_stats);
void ho_target_clear(ho_target* t, const float* rgba, float depth);

#ifdef __cplusplus
}
#endif
#endif
