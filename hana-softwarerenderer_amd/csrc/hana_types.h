// hana_types.h -- data layouts shared by the host runtime and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hana {

enum {
    SHADER_GROUND = 0,
    SHADER_TOON = 1,
    SHADER_TEXTURE = 2,
    SHADER_TEXTURE_LIGHT = 3,
    SHADER_BLINN = 4,
    SHADER_NORMALMAP = 5,
    SHADER_SHADOW = 6
};

// Screen tiles are 2^shift pixels a side: 16x16 (shift 4) by default, 32x32
// (shift 5) where a batch's tuning picks it.  Every kernel that maps pixels
// to tiles is instantiated for both.
constexpr int TILE_SHIFT_DEFAULT = 4;
constexpr int TILE_SHIFT_MAX = 5;
__host__ __device__ constexpr int tile_edge(int shift) { return 1 << shift; }
constexpr int BLOCK = 256;                 // threads per workgroup (4 wave64)
#ifndef HANA_SMALL_PIX
#define HANA_SMALL_PIX 32
#endif
// In passes of over 4 target pixels per face, small triangles of more than
// HANA_MED_PIX candidates are shared by HANA_MED_LANES lanes each instead of
// looped over by one: a wave's trip count is then bounded by HANA_SMALL_PIX /
// HANA_MED_LANES.  Passes of smaller faces
// (micro-triangle meshes) keep one lane per triangle, which their few
// candidates suit better.
#ifndef HANA_MED_PIX
#define HANA_MED_PIX 4
#endif
#ifndef HANA_MED_LANES
#define HANA_MED_LANES 4
#endif
// A raster workgroup (one tile at a time) has 2 or 4 waves: 128 or 256 lanes,
// at most one lane per pixel of a 16x16 tile.
static_assert(HANA_SMALL_PIX >= 1 && HANA_SMALL_PIX <= 32, "small-triangle candidates: at most 32");
#ifndef HANA_VPT
#define HANA_VPT 4
#endif
#ifndef HANA_FPT
#define HANA_FPT 4
#endif
constexpr int VPT = HANA_VPT;              // vertices per thread in k_vertex
constexpr int FPT = HANA_FPT;              // faces per thread in k_setup
// Vertices the device keeps of a clipped polygon: a convex polygon gains at
// most one vertex per plane (3 + 7), rounding in the inside tests can add a
// few more; one lane of a 16-lane clip group holds each.
constexpr int POLY_MAX = 16;
constexpr int FAN_BITS = 4;                // fan index of a clipped triangle in its bin entry
constexpr uint32_t POOL_CAP_MAX = (1u << (31 - FAN_BITS)) - 1u;   // polygons a workspace may hold
constexpr int POLY_VARY = 8;               // varyings stored per clipped vertex
constexpr uint32_t NO_WINNER = 0xFFFFFFFFu;
constexpr uint32_t CLIPPED_BIT = 0x80000000u;

// Uniform block of one pass, derived on the host from ShaderData
// (SR/IShader.h:17-32) with the products the reference recomputes per vertex
// folded once: clip_m = camera_vp*model (ObjectToClipPos, SR/IShader.h:55-57),
// light_m = light_vp*model (ObjectToViewPos, SR/IShader.h:59-61).
struct DevUniforms {
    float clip_m[16];
    float light_m[16];
    float model[16];
    float model_inv[16];
    float light_vp[16];
    float view_pos[4];
    float light_dir[4];
    float light_color[4];
    float ambient[4];
    float mat_color[4];
    float mat_spec[4];
    float gloss;
    float bump_scale;
    int32_t enable_shadow;
    int32_t shadow_fp32;         // fp32 shadow-map extension (hana_uniforms)
};

// One pass = one graphics_draw_triangle call over a mesh into a target.
struct PassDesc {
    // mesh (structure of arrays)
    const float* px; const float* py; const float* pz;
    const float* nx; const float* ny; const float* nz;
    const float* tu; const float* tv;
    const int32_t* idx;          // 3 per face
    int32_t nverts, nfaces;
    int32_t kind, nvary;
    const DevUniforms* u;
    // textures: RGBA8 words in TGAColor byte order
    const uint32_t* diffuse; int32_t dw, dh;
    const uint32_t* normal;  int32_t nw, nh;
    // shadow map sampled by the fragment program (RGBA8 target words, and
    // its fp32 depth for the fp32 extension)
    const uint32_t* shadow; int32_t sw, sh;
    const float* shadow_depth;
    // target
    uint32_t* color; float* depth; uint32_t* winner;
    int32_t width, height, tiles_x, tiles_y;
    int32_t clear;               // 1: the target is cleared to clear_rgba/clear_depth first
    uint32_t clear_rgba; float clear_depth;
    int32_t share_small;         // 1: the pass averages over 4 target pixels per face (see HANA_MED_PIX)
    // workspace slices owned by this pass
    float4* vr;                  // per vertex: ndc.x, ndc.y, ndc.z, 1/w
    float* vs;                   // per vertex: nvary varyings
    uint8_t* code;               // per vertex: bit 7 is_vertex_visible, bit p outside clip plane p
    int32_t* poly_of_face;       // clipped polygon of a face (-1: none)
    uint64_t* stats;             // 5 counters (hana_pass_stats order), or null
    // first block of this pass in each kernel's grid / first global tile
    int32_t vblock0, fblock0, tile0, tblock0;
};

// Buffers shared by every pass of one launch sequence.
struct Workspace {
    uint32_t* tile_count;        // per global tile
    uint32_t* tile_offset;       // exclusive scan of tile_count (+ total at [ntiles])
    uint32_t* tile_cursor;
    uint32_t* bins;              // triangle entries grouped by tile
    float4* pool_vr;             // clipped polygons: POLY_MAX vertices each
    float* pool_vs;              //   ... POLY_VARY varyings per vertex
    int32_t* pool_face;          //   face (within its pass) of each polygon
    int32_t* pool_pass;          //   pass of each polygon
    int32_t* pool_n;             //   vertex count of each polygon
    uint64_t* clip_queue;        // faces failing the trivial accept: pass << 32 | face
    uint4* tris;                 // every triangle with tiles: bin entry, pass, first tile, last tile;
                                 //   BLOCK*FPT per k_setup workgroup, then the clipped fan triangles
    uint32_t* seg_count;         // records in each k_setup workgroup's segment
    uint2* seg_tiles;            // per segment, seg_stride entries: (tile, pairs) of its triangles
                                 //   over at most WIDE_TILES tiles, counted in k_setup's LDS table
    uint16_t* seg_wide;          // per segment, BLOCK*FPT entries: its triangles over more tiles
    uint32_t* seg_info;          // per segment: listed tiles | wide triangles << 16
    uint32_t seg_stride;
    uint64_t* scan_part;         // per scan chunk: sum of its tile counts
    uint32_t* status;            // see STATUS_*
    uint32_t* status_out;        // or null: host-mapped copy of the status words, written by the scan
    int32_t ntiles;
    uint32_t bin_cap;
    uint32_t poly_cap;
    uint32_t clipq_cap;
    uint32_t tris_cap;
    uint32_t seg_region;         // records reserved for the setup segments
    // queue of the persistent raster (k_raster_q): its items are the tiles of
    // every pass in global tile order (shadow passes first, then main passes
    // that sample no map drawn here, then those that do)
    uint32_t* qstate;            // [0] next item, [1] group-0 tiles finished and published
    uint32_t q_items;            // tiles of all passes
    uint32_t q_g0;               // tiles of group-0 passes
    uint32_t q_dep_from;         // first item of a pass that samples a map drawn by group 0
    int32_t tile_shift;          // tiles are 2^tile_shift pixels a side
    int32_t raster_waves;        // waves per raster workgroup: 2 or 4
    // per raster group, or null: for every run of 2^TILE_PASS_SHIFT
    // workgroups of k_raster, the pass (from the group's first) of the run's
    // first tile, with TILE_PASS_PURE set when the whole run lies in it, and
    // that pass's tile0 - tblock0 (workgroup b of it renders global tile
    // b + that)
    const uint2* tile_pass[2];
};
constexpr int TILE_PASS_SHIFT = 6;
constexpr uint32_t TILE_PASS_PURE = 0x80000000u;

// the tile scan: SCAN_CHUNK tile counters per workgroup of SCAN_THREADS
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER_THREAD = 16;
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_PER_THREAD;

// status words.  The first STATUS_SEQ_WORDS belong to one launch sequence
// (zeroed by its first kernel): polygons allocated, faces queued for
// clipping, clipped fan triangles listed, bin entries needed, and the flags
// this sequence raised.  The others stay set across sequences until the host
// reads them at a sync point: the flags of every sequence and the largest
// needs seen, so that calls need not wait for one another.
enum { STATUS_POLYS = 0, STATUS_CLIPQ = 1, STATUS_TRIS = 2, STATUS_BINS_NEEDED = 3, STATUS_SEQ_FLAGS = 4,
       STATUS_SEQ_WORDS = 5, STATUS_FLAGS = 5, STATUS_NEED_BINS = 6, STATUS_NEED_POLYS = 7, STATUS_NEED_TRIS = 8,
       STATUS_WORDS = 9 };
enum { FLAG_BIN_OVERFLOW = 1, FLAG_POLY_OVERFLOW = 2, FLAG_POLY_TOO_LARGE = 4, FLAG_CLIPQ_OVERFLOW = 8,
       FLAG_TRIS_OVERFLOW = 16, FLAG_PAIR_LIMIT = 32, FLAG_QUEUE_STALL = 64 };
// A workspace capacity ran out: the sequence is incomplete, its rasters write
// nothing, and the host grows the workspace and renders it again.
constexpr uint32_t FLAG_CAPACITY = FLAG_BIN_OVERFLOW | FLAG_POLY_OVERFLOW | FLAG_CLIPQ_OVERFLOW | FLAG_TRIS_OVERFLOW;
// Flags of a sequence whose binning and rasters do nothing.
constexpr uint32_t FLAG_NO_RASTER = FLAG_CAPACITY | FLAG_PAIR_LIMIT;
// (triangle, tile) pairs one launch sequence may hold: bin positions are
// 32-bit.  A sequence with more sets FLAG_PAIR_LIMIT and rasterises nothing.
constexpr uint32_t MAX_PAIRS = 0xFFFFFFF0u;
// vertex outcode bits
constexpr uint32_t CODE_VISIBLE = 0x80u;
constexpr uint32_t CODE_PLANES = 0x7Fu;
enum { ST_FACES = 0, ST_CLIPPED = 1, ST_POLYS = 2, ST_ENTRIES = 3, ST_PIXELS = 4, ST_WORDS = 5 };

}  // namespace hana
