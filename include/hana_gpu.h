/*
 * hana_gpu.h -- C ABI of the MI355X (gfx950) rasterisation pipeline.
 *
 * Plain C types only: pointers, sizes, PODs.  Every function returns 0 on
 * success and a negative HANA_E* code on failure; hana_gpu_last_error()
 * describes the last failure on a context.  No exception crosses this ABI.
 *
 * Each entry point names the reference interface it replaces
 * (SR/x:N = Hana-SoftwareRenderer/x line N of the reference):
 *
 *   hana_draw            graphics_draw_triangle(DrawData*)         SR/graphics.h:15, SR/graphics.cpp:378-407
 *   hana_render_frame    DrawModel::draw(Camera*, RenderBuffer*, bool)  SR/scene.h:53-99
 *   hana_mesh_create     Model accessors vert/normal/uv             SR/model.h:19-33, SR/model.cpp:66-111
 *   hana_texture_create  TGAImage::get / TGAColor(p, bpp)           SR/tgaimage.h:46-53,66-98, SR/tgaimage.cpp:248-253
 *   hana_target_*        RenderBuffer ctor / clears / buffers       SR/renderbuffer.h:5-21, SR/renderbuffer.cpp:3-76
 *   hana_uniforms        ShaderData + Material (no pointers)        SR/IShader.h:7-32
 *   hana_build_uniforms  DrawModel::draw uniform set-up             SR/scene.h:55-71, SR/camera.cpp:82-92
 *   hana_target_present  window_draw_buffer (BGR swap + row flip)   SR/win32.cpp:348-370
 *   hana_obj_load        Model::Model OBJ parser + corner fetches   SR/model.cpp:6-48,66-111
 *   hana_tga_load/save   TGAImage::read_tga_file / write_tga_file   SR/tgaimage.cpp:40-246
 *   HANA_SHADER_*        the IShader subclasses                     SR/IShader.h:132-165
 *   hana_scene_*         Scene / SingleModelScene state, orbit      SR/scene.cpp:3-9,45-63,127-147,
 *                        camera and its reset                       SR/camera.cpp:31-70,167-169
 *
 * A context owns four HIP streams, each with its own workspace, and every
 * device object created on it; calls on one context must be externally
 * synchronised.  Work is enqueued asynchronously: consecutive draws rotate
 * over the streams (over two of them for draws of 2^19 faces or more) so that
 * a draw's geometry stages overlap the rasters of the draws before it.  A
 * raster waits for every earlier draw that draws or samples one of its
 * targets or shadow maps, so each target sees the draws in call order; draws
 * on disjoint targets may run concurrently.  hana_target_read(_winner),
 * hana_target_present and hana_gpu_sync wait for all of it.  A draw that
 * outgrows the workspace (bins, clip pool) leaves its targets untouched; the
 * next such wait grows the workspace and renders the draws since the previous
 * wait again, in call order, before it returns (the bin overflow of
 * SR/graphics.cpp has no counterpart: the reference draws without a bound).
 * Uploads, batch runs and the destruction of objects a pending draw uses wait
 * the same way first.
 */
#ifndef HANA_GPU_H
#define HANA_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HANA_OK = 0,
    HANA_E_INVALID = -1,     /* bad argument */
    HANA_E_DEVICE = -2,      /* HIP runtime error */
    HANA_E_NOMEM = -3,       /* device allocation failed or over budget */
    HANA_E_OVERFLOW = -4,    /* internal capacity exceeded after three replays */
    HANA_E_UNSUPPORTED = -5  /* a clipped polygon had more than 16 vertices: it was
                                dropped, the rest drawn (the reference keeps 30) */
};

/* IShader subclasses (SR/IShader.h:132-165). */
enum {
    HANA_SHADER_GROUND = 0,
    HANA_SHADER_TOON = 1,
    HANA_SHADER_TEXTURE = 2,
    HANA_SHADER_TEXTURE_LIGHT = 3,
    HANA_SHADER_BLINN = 4,
    HANA_SHADER_NORMALMAP = 5,
    HANA_SHADER_SHADOW = 6
};

typedef struct hana_ctx hana_ctx;
typedef struct hana_mesh hana_mesh;
typedef struct hana_texture hana_texture;
typedef struct hana_target hana_target;
typedef struct hana_batch hana_batch;

/* ShaderData + Material scalars (SR/IShader.h:7-32); row-major 4x4. */
typedef struct hana_uniforms {
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
    /* 0: the reference's shadow map (parity mode): the shadow pass writes
     * White * clamp01(light clip z) as RGBA8 and is_in_shadow compares
     * against R/255 (SR/IShader.cpp:176-180, SR/IShader.h:124).
     * 1: fp32 extension (no reference counterpart): the shadow pass writes
     * depth only, and is_in_shadow compares against 2*d - 1, the light clip z
     * of the nearest caster from the map's fp32 depth d (an uncovered texel
     * holds the clear depth and lights).  Set it alike for a frame's shadow
     * and main passes. */
    int32_t shadow_fp32;
} hana_uniforms;

/* One graphics_draw_triangle call (DrawData, SR/graphics.h:9-13). */
typedef struct hana_draw_desc {
    hana_mesh* mesh;
    int32_t shader;                 /* HANA_SHADER_* */
    int32_t pad_;
    const hana_uniforms* uniforms;
    hana_texture* diffuse;          /* Material::diffuse_map, may be NULL */
    hana_texture* normal;           /* Material::normal_map, may be NULL */
    hana_target* shadow_map;        /* ShaderData::shadow_map, may be NULL */
    hana_target* target;            /* DrawData::render_buffer */
} hana_draw_desc;

/* One DrawModel::draw call (SR/scene.h:53-99). */
typedef struct hana_frame_desc {
    hana_mesh* mesh;
    int32_t shader;
    int32_t enable_shadow;
    const hana_uniforms* uniforms;
    hana_texture* diffuse;
    hana_texture* normal;
    hana_target* shadow_map;        /* required when enable_shadow */
    hana_target* target;
} hana_frame_desc;

/* Per-pass counters (filled when stats are enabled on the context). */
typedef struct hana_pass_stats {
    int64_t faces;           /* faces submitted */
    int64_t clipped_faces;   /* faces failing the trivial-accept test */
    int64_t polygons;        /* clipped polygons with >= 3 vertices */
    int64_t tile_entries;    /* (triangle, tile) bin entries */
    int64_t pixels_written;  /* pixels whose colour/depth the pass wrote */
} hana_pass_stats;

/* ---- context ---- */
int hana_gpu_create(int device, hana_ctx** out);
int hana_gpu_destroy(hana_ctx* ctx);
const char* hana_gpu_last_error(hana_ctx* ctx);
int hana_gpu_sync(hana_ctx* ctx);
/* Collect hana_pass_stats for subsequent draws (adds a few atomics). */
int hana_gpu_set_stats(hana_ctx* ctx, int enable);
int hana_gpu_last_stats(hana_ctx* ctx, hana_pass_stats* out, int max_passes, int* n_passes);
/* How many times a wait found a draw that outgrew the workspace and rendered
 * the draws since the previous wait again (no counterpart in the reference,
 * whose face loop has no capacity bound). */
int hana_gpu_replay_count(hana_ctx* ctx, int64_t* out);

/* Per-stage kernel timing with HIP events recorded around each launch
 * (benchmarking; while it is on, draws use one stream and do not overlap).  Stages: 0 vertex, 1 set-up (cull,
 * bounds, tile counts), 2 clipping, 3 tile scan, 4 binning, 5 (unused, no
 * launches), 6 raster of the first dependency group (shadow passes), 7 raster
 * of the second group (main passes).  Reading returns the milliseconds and
 * launch counts accumulated since the last read. */
#define HANA_TIMING_STAGES 8
int hana_gpu_set_timing(hana_ctx* ctx, int enable);
int hana_gpu_read_timing(hana_ctx* ctx, double* ms, int64_t* launches);

/* ---- resources ---- */
/* SoA fp32 attributes (pos 3V, nrm 3V, uv 2V) and int32 indices (3F); one
 * index per corner for position, normal and uv.  Normals are used as
 * Model::normal returns them: renormalised (SR/model.cpp:108-111). */
int hana_mesh_create(hana_ctx* ctx, int32_t nverts, const float* pos, const float* nrm,
                     const float* uv, int32_t nfaces, const int32_t* idx, hana_mesh** out);
int hana_mesh_destroy(hana_ctx* ctx, hana_mesh* m);
/* TGAImage bytes: bpp in {1,3,4}, B,G,R,A order, texel (x,y) at (x+y*w)*bpp. */
int hana_texture_create(hana_ctx* ctx, int32_t w, int32_t h, int32_t bpp, const uint8_t* bytes,
                        hana_texture** out);
int hana_texture_destroy(hana_ctx* ctx, hana_texture* t);
/* RenderBuffer(width, height): RGBA8 at (y*W+x)*4 and fp32 depth at y*W+x,
 * initialised like the reference constructor (colour 0,0,0,255, depth 1).
 * 1 <= w, h <= 2^20 and w*h <= 2^30, else HANA_E_INVALID. */
int hana_target_create(hana_ctx* ctx, int32_t w, int32_t h, hana_target** out);
int hana_target_destroy(hana_ctx* ctx, hana_target* t);
/* renderbuffer_clear_color + renderbuffer_clear_depth; deferred and fused
 * into the next pass that draws into the target. rgba bytes as written. */
int hana_target_clear(hana_ctx* ctx, hana_target* t, const uint8_t rgba[4], float depth);
int hana_target_upload(hana_ctx* ctx, hana_target* t, const uint8_t* rgba, const float* depth);
int hana_target_read(hana_ctx* ctx, hana_target* t, uint8_t* rgba, float* depth);
/* Optional per-pixel winner map: for each written pixel the primitive id
 * face*16 + (trivially accepted ? 0 : 1 + fan index), 0xFFFFFFFF elsewhere.
 * Enabled until the target is destroyed; read with hana_target_read_winner. */
int hana_target_enable_winner(hana_ctx* ctx, hana_target* t, int enable);
int hana_target_read_winner(hana_ctx* ctx, hana_target* t, uint32_t* winner);

/* window_draw_buffer without a window (SR/win32.cpp:348-370): the target's
 * colour as a top-down BGRA8 image, row r of the target at row H-1-r, bytes
 * B,G,R,A (alpha carried from the target).  bgra holds W*H*4 bytes; waits. */
int hana_target_present(hana_ctx* ctx, hana_target* t, uint8_t* bgra);

/* ---- drawing ---- */
int hana_draw(hana_ctx* ctx, const hana_draw_desc* d);
int hana_render_frame(hana_ctx* ctx, const hana_frame_desc* f);
/* hana_render_frame for callers that cannot allocate (code under stream
 * capture, serving loops that must not call hipMalloc); no counterpart in the
 * reference.  The same frame and result, on the context stream, with the
 * workspace in a fixed device array of 128 MiB per GPU (held until the
 * process ends) instead of allocated buffers: no hipMalloc, hipFree or other
 * allocating call.  A frame that needs more workspace is rejected with
 * HANA_E_NOMEM before anything runs; one whose bins or clip pool overflow at
 * run time leaves its targets untouched and the next wait returns
 * HANA_E_OVERFLOW.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (two contexts, two threads) race.  Refused
 * (HANA_E_INVALID) while per-call draws are pending (hana_gpu_sync first) or
 * with stage timing or statistics on. */
int hana_render_frame_noalloc(hana_ctx* ctx, const hana_frame_desc* f);

/* ---- tuning: how launch sequences are laid out ----
 * No reference counterpart (the reference scans each triangle's bounding box
 * on one core).  tile_size: screen tiles of 16x16 or 32x32 pixels (binning and
 * raster).  raster_wgs_per_cu: 0 rasterises each pass group with one
 * workgroup per tile; n > 0 runs one persistent raster of n workgroups per CU
 * that takes the tiles of every pass from a queue (shadow passes first, main
 * passes waiting for them without a second launch).  raster_waves: lanes per
 * raster workgroup, 2 or 4 waves (two waves fit twice the tiles on a CU).
 * parts: the concurrent parts a batch is split into (1-4; ignored for
 * per-call draws).  Results are identical under every tuning.  Default: 16,
 * 0, 4 parts, 4 waves. */
typedef struct hana_tuning {
    int32_t tile_size;
    int32_t raster_wgs_per_cu;
    int32_t parts;
    int32_t raster_waves;           /* waves per raster workgroup: 2 or 4 (0: 4) */
    double ms;                      /* hana_batch_autotune: measured run time of the choice */
} hana_tuning;
/* Tuning of the context's per-call draws (hana_draw, hana_render_frame). */
int hana_gpu_set_tuning(hana_ctx* ctx, const hana_tuning* t);

/* ---- batches: many frames in one captured graph ----
 * The frames are split into up to four parts that run concurrently (branches
 * of the graph on the context's streams and the batch's own); they need
 * distinct targets and shadow maps.  hana_batch_run returns once the work is
 * enqueued; it first completes the per-call draws enqueued before it. */
int hana_batch_create(hana_ctx* ctx, hana_batch** out);
int hana_batch_add_frame(hana_ctx* ctx, hana_batch* b, const hana_frame_desc* f);
/* Tuning used by hana_batch_finalize (before it). */
int hana_batch_set_tuning(hana_ctx* ctx, hana_batch* b, const hana_tuning* t);
/* Build the workspace and (when use_graph) capture the launch sequence. */
int hana_batch_finalize(hana_ctx* ctx, hana_batch* b, int use_graph);
/* Finalizes the batch with the tuning that ran it fastest: from the batch's
 * tuning, tile size, then raster form (per tile, or 2, 4 or 8 persistent
 * workgroups per CU), then raster waves (4, 2), then parts (1, 2, 4), one
 * dimension at a time, each
 * candidate sized, captured and timed over `reps` runs; a candidate must be
 * more than 2 % faster than the best so far to replace it.  The frames are
 * rendered by every run.  `chosen` (nullable) receives the choice and its
 * time. */
int hana_batch_autotune(hana_ctx* ctx, hana_batch* b, int use_graph, int reps, hana_tuning* chosen);
int hana_batch_get_tuning(hana_ctx* ctx, const hana_batch* b, hana_tuning* out);
int hana_batch_run(hana_ctx* ctx, hana_batch* b);
int hana_batch_destroy(hana_ctx* ctx, hana_batch* b);

/* ---- serving: continuous batching of render requests ----
 * No single reference counterpart: the reference renders one scene per
 * window frame (SR/main.cpp:87-156, DrawModel::draw SR/scene.h:53-99).  A
 * server keeps meshes and textures resident, admits requests while the
 * in-flight HBM footprint stays within hbm_budget_bytes, forms a batch of
 * whatever is admitted whenever one of its lanes (HIP streams) is idle, and
 * runs lanes concurrently.  Every request is a DrawModel::draw frame on
 * cleared targets; its colour image is copied back before completion. */
typedef struct hana_server hana_server;

enum { HANA_SERVER_RETURN_DEPTH = 1 };   /* completions carry the fp32 depth buffer too */

typedef struct hana_server_config {
    int32_t lanes;                /* concurrent batches (streams), >= 1 */
    int32_t max_batch_frames;     /* frames per batch */
    int64_t max_batch_pixels;     /* frame + shadow pixels per batch */
    int64_t hbm_budget_bytes;     /* in-flight device footprint limit */
    int32_t flags;                /* HANA_SERVER_* */
    int32_t pad_;
} hana_server_config;

typedef struct hana_request {
    hana_mesh* mesh;
    hana_texture* diffuse;
    hana_texture* normal;
    int32_t shader;               /* HANA_SHADER_* */
    int32_t enable_shadow;
    int32_t width, height;
    int32_t shadow_width, shadow_height;   /* 0: frame-sized, as the reference */
    float cam_pos[3];
    float cam_target[3];
    float light_pos[3];
    int32_t shadow_fp32;          /* hana_uniforms::shadow_fp32 */
    uint64_t user_tag;
} hana_request;

typedef struct hana_completion {
    uint64_t id;
    uint64_t user_tag;
    int32_t status;               /* HANA_OK or a negative code */
    int32_t width, height;
    int32_t batch_frames;         /* size of the batch it ran in */
    const uint8_t* rgba;          /* W*H*4 image, valid until hana_server_release */
    double t_submit_ms, t_launch_ms, t_done_ms;   /* server clock */
    const float* depth;           /* W*H depth (HANA_SERVER_RETURN_DEPTH), else NULL; same lifetime */
} hana_completion;

typedef struct hana_server_stats {
    int64_t submitted, completed, rejected, batches, retries;
    int64_t inflight_bytes, peak_inflight_bytes;
    double busy_ms;               /* time with at least one batch in flight */
} hana_server_stats;

int hana_server_create(hana_ctx* ctx, const hana_server_config* cfg, hana_server** out);
/* Thread-safe; returns the request id. */
int hana_server_submit(hana_server* s, const hana_request* r, uint64_t* id);
/* Up to max completions; waits up to wait_ms for the first one. */
int hana_server_poll(hana_server* s, hana_completion* out, int max, int* n, int wait_ms);
int hana_server_release(hana_server* s, uint64_t id);
int hana_server_get_stats(hana_server* s, hana_server_stats* out);
double hana_server_now_ms(hana_server* s);
int hana_server_destroy(hana_server* s);

/* ---- on-disk formats (host only, no context) ----
 * Errors return HANA_E_INVALID / HANA_E_NOMEM with hana_asset_last_error()
 * describing them (per thread). */
const char* hana_asset_last_error(void);

/* A Wavefront OBJ as Model reads it (SR/model.cpp:6-48: "v", "vt", "vn" and
 * "f v/vt/vn ..." lines, 1-based, the first three corners of a face used).
 * Normals are the values graphics_draw_triangle's first pass over a freshly
 * loaded Model fetches: Model::normal renormalises the stored vector at every
 * fetch (SR/model.cpp:108-111), corners in face order.  One vertex per
 * distinct (position, uv, fetched normal); arrays malloc'ed, freed by
 * hana_obj_free.  Faces with fewer than three v/vt/vn corners, or indexing
 * past a list, are rejected (the reference reads out of bounds there). */
typedef struct hana_obj_mesh {
    int32_t nverts, nfaces;
    float* pos;                     /* 3*nverts */
    float* nrm;                     /* 3*nverts */
    float* uv;                      /* 2*nverts */
    int32_t* idx;                   /* 3*nfaces */
} hana_obj_mesh;
int hana_obj_load(const char* path, hana_obj_mesh* out);
int hana_obj_free(hana_obj_mesh* m);

/* TGAImage memory layout: bytespp in {1,3,4}, B,G,R(,A) bytes, texel (x,y)
 * at (x + y*width)*bytespp. */
typedef struct hana_image {
    int32_t width, height, bytespp, pad_;
    uint8_t* bytes;
} hana_image;
/* TGAImage::read_tga_file (SR/tgaimage.cpp:40-143): raw (2, 3) or RLE
 * (10, 11) data read right after the 18-byte header; flipped to top-down
 * unless the descriptor says so, mirrored if right-to-left.  flip_vertically
 * adds the extra flip of Model::load_texture (SR/model.cpp:80-81).  A
 * malformed RLE stream is an error (the reference writes past its buffer). */
int hana_tga_load(const char* path, int flip_vertically, hana_image* out);
/* TGAImage::write_tga_file (SR/tgaimage.cpp:145-246): top-down header, raw or
 * RLE packets as the reference forms them, TGA 2.0 footer; byte-identical. */
int hana_tga_save(const char* path, const hana_image* img, int rle);
int hana_image_free(hana_image* img);

/* ---- interactive scene state ----
 * The state the reference's main loop keeps between frames: the orbit camera,
 * the light, the shader and shadow switches.  A sequence of inputs applied
 * here, then hana_scene_uniforms + hana_render_frame per frame, renders what
 * SingleModelScene::tick draws for the same inputs. */
typedef struct hana_motion {   /* Motion (SR/camera.h:11): deltas in window heights */
    float orbit[2];
    float pan[2];
    float dolly;
} hana_motion;

typedef struct hana_scene {
    float camera_position[3];
    float camera_target[3];   /* moved by panning; lights and shadows use the origin (SR/camera.cpp:94-97) */
    float aspect;
    float light_position[3];
    int32_t shader_index;     /* 0 NormalMap, 1 Blinn, 2 Texture+Light, 3 Texture, 4 Toon, 5 Ground (SR/scene.cpp:86-97) */
    int32_t enable_shadow;
} hana_scene;

enum {   /* the keys the scene reacts to (SR/platform.h:7) */
    HANA_KEY_A = 0,      /* light rotates -5 degrees about (0,1,1) */
    HANA_KEY_D = 1,      /* light rotates +5 degrees */
    HANA_KEY_E = 5,      /* shadow on/off */
    HANA_KEY_Q = 4,      /* next shader; shadow on for the first three */
    HANA_KEY_SPACE = 6   /* camera back to (0,0,2) -> origin */
};

/* Scene::Scene for a width x height framebuffer: camera (0,0,2) -> origin,
 * light (2,2,2), shadow on, NormalMap. */
int hana_scene_init(hana_scene* s, int32_t width, int32_t height);
/* Camera::update_transform (pan, then orbit and dolly about the target). */
int hana_scene_update_camera(hana_scene* s, const hana_motion* m);
/* Scene::on_key_input / SingleModelScene::on_key_input for a pressed key. */
int hana_scene_key(hana_scene* s, int32_t key);
/* The HANA_SHADER_* kind of the scene's current shader. */
int hana_scene_shader(const hana_scene* s);
/* hana_build_uniforms for the scene's camera, light and shadow switch. */
int hana_scene_uniforms(const hana_scene* s, const float obj_position[3], const float obj_rotation_deg[3],
                        const float obj_scale[3], int32_t fb_w, int32_t fb_h, hana_uniforms* out);

/* ---- host uniform math ---- */
/* ShaderData as DrawModel::draw fills it for Camera(cam_pos, cam_target,
 * cam_aspect), a light GameObject at light_pos, the model transform
 * (position, rotation in degrees, scale) and a framebuffer fb_w x fb_h. */
int hana_build_uniforms(const float cam_pos[3], const float cam_target[3], float cam_aspect,
                        const float light_pos[3], const float obj_position[3],
                        const float obj_rotation_deg[3], const float obj_scale[3], int32_t fb_w,
                        int32_t fb_h, int32_t enable_shadow, hana_uniforms* out);

#ifdef __cplusplus
}
#endif
#endif
