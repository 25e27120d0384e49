// hana_internal.h -- runtime objects shared by the C ABI implementation
// files (hana_runtime.cpp, hana_server.cpp).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/hana_gpu.h"
#include "hana_types.h"

struct hana_mesh {
    int32_t V = 0, F = 0;
    float* soa = nullptr;      // px | py | pz | nx | ny | nz | u | v, V floats each
    int32_t* idx = nullptr;    // 3F
};

struct hana_texture {
    int32_t w = 0, h = 0;
    uint32_t* texels = nullptr;   // TGAColor bytes b,g,r,a per word
};

struct hana_target {
    int32_t w = 0, h = 0;
    uint32_t* color = nullptr;    // RGBA8 (R in the low byte)
    float* depth = nullptr;
    uint32_t* winner = nullptr;
    bool pending_clear = false;
    uint32_t clear_rgba = 0;
    float clear_depth = 0;
};

namespace hana {
namespace rt {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct PassSpec {
    hana_mesh* mesh;
    int kind;
    hana_uniforms u;
    hana_texture* diffuse;
    hana_texture* normal;
    hana_target* shadow;    // sampled by the fragment program
    hana_target* target;
    int group;              // raster dependency group (shadow passes first)
};

// How a launch sequence is laid out: tile size, raster form, batch parts
// (hana_tuning of the C ABI).
struct Tuning {
    int tile_shift = TILE_SHIFT_DEFAULT;   // tiles of 2^shift pixels a side
    int raster_wgs_per_cu = 0;            // 0: one raster workgroup per tile; else a persistent raster
    int raster_waves = 4;                  // waves per raster workgroup (2 or 4)
    int parts = 4;                         // concurrent parts of a batch
};

// Grow-only device workspace for launch sequences.
struct Arena {
    // tables: the pass descriptors, then (at Plan::uni_off) their uniforms,
    // then (at Plan::tp_off) the raster's pass lookup, uploaded by one copy
    // per call
    DevBuf tables, vr, vs, code, pof, tile_count, tile_offset, tile_cursor, bins, tris, seg_count, scan_part;
    DevBuf seg_tiles, seg_wide, seg_info, qstate;
    DevBuf pool_vr, pool_vs, pool_face, pool_pass, pool_n, clip_queue, status, stats;
    uint32_t bin_cap = 0, poly_cap = 0, clipq_cap = 0, tris_cap = 0;
};

struct Plan {
    std::vector<PassSpec> specs;
    std::vector<PassDesc> descs;
    std::vector<DevUniforms> unis;
    // per-pass workspace slice offsets (vertices, varying floats, faces)
    std::vector<size_t> v_off, vs_off, f_off;
    int vblocks = 0, fblocks = 0;
    int ngroups = 0;
    int group_first[2] = {0, 0}, group_count[2] = {0, 0}, group_blocks[2] = {0, 0};
    size_t total_v = 0, total_vs = 0, total_f = 0;
    size_t uni_off = 0;   // byte offset of the uniforms in the tables buffer
    // k_raster's pass lookup (Workspace::tile_pass) of both groups, at byte
    // tp_off of the tables buffer; group g's entries start at tp_first[g]
    std::vector<uint2> tile_pass;
    int tp_first[2] = {0, 0};
    size_t tp_off = 0;
    int total_tiles = 0;
    bool small_tris = true;   // every pass has at most 4 target pixels per face
    int tile_shift = TILE_SHIFT_DEFAULT;
    int raster_waves = 4;
    int raster_wgs = 0;       // workgroups of the persistent raster (0: one per tile and group)
    uint32_t q_g0 = 0, q_dep_from = 0;
    Workspace ws{};
};

// A clear fused into nothing: a shadow map sampled by a call while a clear of
// it was pending is filled with the clear values before that call's raster.
struct Fill {
    hana_target* target;
    uint32_t rgba;
    float depth;
};

// A per-call launch sequence as enqueued, kept until the next status check
// so that it can be rendered again if a workspace capacity ran out.
struct LoggedCall {
    Plan pl;
    std::vector<Fill> fills;
};

// Size of the device array that holds the workspace of
// hana_render_frame_noalloc (one per GPU, for the life of the process).
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;

constexpr int kRing = 32;
// calls kept for replay; past this many a call first checks the status of
// those before it (a sync)
constexpr size_t kMaxLoggedCalls = 1024;
// streams (each with a workspace) that consecutive per-call draws rotate over
#ifndef HANA_DRAW_SLOTS
#define HANA_DRAW_SLOTS 4
#endif
constexpr int kDrawSlots = HANA_DRAW_SLOTS;
// launch sequences of at least this many faces use two of the slots
constexpr size_t kLargeSequenceFaces = size_t(1) << 19;
// a draw waits for another slot once that slot has more targets pending for
// it than this, conflicting or not, which bounds the bookkeeping
constexpr size_t kMaxPendingTargets = 1024;
static_assert(kDrawSlots >= 1, "a context has at least its own stream");

}  // namespace rt
}  // namespace hana

struct hana_ctx {
    using Arena = hana::rt::Arena;
    int device = 0;
    int num_cus = 256;
    hana::rt::Tuning tuning;   // of per-call draws
    int nslots = 1;   // draw slots of this context (at most kDrawSlots)
    hipStream_t stream = nullptr;
    std::string err;
    bool stats_enabled = false;
    Arena arena;
    // pinned staging ring for per-call descriptor uploads
    void* ring_host[hana::rt::kRing] = {};
    hipEvent_t ring_event[hana::rt::kRing] = {};
    bool ring_used[hana::rt::kRing] = {};
    size_t ring_bytes = 0;
    int ring_next = 0;
    // status of the last launch sequence
    uint32_t* status_host = nullptr;   // pinned, STATUS_WORDS
    uint64_t* stats_host = nullptr;    // pinned, per pass
    int stats_passes = 0;
    bool status_pending = false;
    // pipelined draws: consecutive draws rotate over nslots slots, each a
    // stream with its own workspace and status words (slot 0 is (stream,
    // arena, status_host), slot k > 0 is (xstream, xarena, xstatus_host)[k-1]),
    // so the geometry stages of a draw overlap the rasters of the draws before;
    // each raster waits for the whole previous draw
    hipStream_t xstream[hana::rt::kDrawSlots - 1] = {};
    Arena xarena[hana::rt::kDrawSlots - 1];
    uint32_t* xstatus_host[hana::rt::kDrawSlots - 1] = {};
    uint32_t* status_dev[hana::rt::kDrawSlots] = {};   // device addresses of the slots' status words
    hipEvent_t slot_done[hana::rt::kDrawSlots] = {};
    // pend[k][b]: the targets (drawn or sampled) of draws on slot k that slot
    // b's stream has not waited for; a draw on b waits for slot k only when
    // it touches one of them
    std::vector<hana_target*> pend[hana::rt::kDrawSlots][hana::rt::kDrawSlots];
    int next_slot = 0;
    bool pipelined = true;
    // device staging for hana_target_present
    hana::rt::DevBuf present;
    // optional per-stage timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    // per-call sequences since the last status check, in call order; a check
    // that finds a capacity overflow grows the workspaces and renders them
    // again (replay_overflow), else reports HANA_E_OVERFLOW
    std::vector<hana::rt::LoggedCall> log;
    bool replay_overflow = true;
    int64_t replays = 0;   // status checks that rendered the logged calls again
    // hana_render_frame_noalloc: the device address of this GPU's scratch
    // array, host-mapped status words of its sequences, whether one ran
    // since the last status check, and the draw slots that must wait for it
    void* scratch = nullptr;
    uint32_t* noalloc_status_host = nullptr;
    uint32_t* noalloc_status_dev = nullptr;
    bool noalloc_pending = false;
    uint32_t noalloc_wait_mask = 0;
};

// A batch's frames are split into up to kMaxBatchParts parts, each a launch
// sequence with its own workspace; the parts run concurrently, part p on the
// stream of draw slot p and parts beyond the slots on streams of the batch
// (branches of one graph), since their frames share no target.
constexpr size_t kMaxBatchParts = 4;

struct hana_batch_part {
    std::vector<hana::rt::PassSpec> specs;
    hana::rt::Plan plan;
    hana::rt::Arena* arena = nullptr;   // the batch's arena of the same index
    size_t f0 = 0, f1 = 0;               // its frames [f0, f1) of the batch
};

struct hana_batch {
    std::vector<hana::rt::PassSpec> specs;
    std::vector<size_t> frame_first;   // index in specs of each frame's first pass
    std::vector<hana_batch_part> parts;
    hana::rt::Arena arenas[kMaxBatchParts];   // grow-only, kept across tunings
    hana::rt::Tuning tuning;
    double tuned_ms = 0.0;                    // run time the autotuner measured for `tuning`
    bool finalized = false;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joined;    // one per part after the first
    std::vector<hipStream_t> extra;    // streams of the parts beyond the context's draw slots
};

namespace hana {
namespace rt {

int fail(hana_ctx* c, int code, const std::string& msg);
int hip_fail(hana_ctx* c, hipError_t e, const char* what);
int ensure(hana_ctx* c, DevBuf& b, size_t bytes);
int materialize_clear(hana_ctx* c, hana_target* t, hipStream_t s);
// A context with nslots draw slots (1..kDrawSlots); hana_gpu_create makes kDrawSlots.
int create_context(int device, int nslots, hana_ctx** out);
// Orders the context stream after every draw in flight on another slot's stream.
int join(hana_ctx* c);
// Waits for every stream of the context.
int sync_all(hana_ctx* c);
// Draw slot k's stream, workspace and host status words (slot 0: the context's own).
hipStream_t slot_stream(hana_ctx* c, int k);
Arena& slot_arena(hana_ctx* c, int k);
uint32_t* slot_status(hana_ctx* c, int k);
int check_status(hana_ctx* c);
// Enqueues the launch sequence of a set of passes on the context's stream.
int run_passes(hana_ctx* c, std::vector<PassSpec> specs);
// Validates a hana_tuning into a Tuning.
int tuning_from(hana_ctx* c, const hana_tuning* t, Tuning& out);
// Resolves the device address of the no-alloc scratch array of the context's
// (current) device, once per device (hana_noalloc.cpp).
int noalloc_init(hana_ctx* c);
bool valid_kind(int k);

}  // namespace rt
}  // namespace hana
