// hana_runtime.cpp -- the C ABI (include/hana_gpu.h): device objects, pass
// planning, workspace management and the kernel launch sequence.
//
// A pass is one graphics_draw_triangle call (SR/graphics.cpp:378-407).  A
// frame is DrawModel::draw (SR/scene.h:53-99): an optional shadow pass into
// the shadow map, the main pass sampling it, then the shadow-map clear.  Any
// set of passes is rendered by one launch sequence of five kernels over
// per-pass descriptor tables (vertex, setup-count, scan, setup-scatter,
// raster per dependency group), so a single frame and a batch of hundreds of
// frames run the same code; batches are captured once into a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hana_gpu.h"
#include "hana_hostmath.h"
#include "hana_internal.h"
#include "hana_launch.h"
#include "hana_types.h"

using namespace hana;
using namespace hana::rt;

namespace hana {
namespace rt {

int fail(hana_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}
int hip_fail(hana_ctx* c, hipError_t e, const char* what) {
    return fail(c, HANA_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIP_TRY(c, expr)                                   \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_fail(c, e_, #expr); \
    } while (0)

hipStream_t slot_stream(hana_ctx* c, int k) { return k ? c->xstream[k - 1] : c->stream; }
Arena& slot_arena(hana_ctx* c, int k) { return k ? c->xarena[k - 1] : c->arena; }
uint32_t* slot_status(hana_ctx* c, int k) { return k ? c->xstatus_host[k - 1] : c->status_host; }

int sync_all(hana_ctx* c) {
    for (int k = 0; k < c->nslots; k++)
        if (slot_stream(c, k)) HIP_TRY(c, hipStreamSynchronize(slot_stream(c, k)));
    for (auto& row : c->pend)
        for (auto& p : row) p.clear();   // every draw has completed
    return HANA_OK;
}

int join(hana_ctx* c) {
    for (int k = 1; k < c->nslots; k++)
        if (!c->pend[k][0].empty()) {
            HIP_TRY(c, hipStreamWaitEvent(c->stream, c->slot_done[k], 0));
            c->pend[k][0].clear();
        }
    return HANA_OK;
}

// The distinct targets a set of passes draws or samples.
static std::vector<hana_target*> touched_targets(const std::vector<PassSpec>& specs) {
    std::vector<hana_target*> t;
    for (const PassSpec& s : specs)
        for (hana_target* x : {s.target, s.shadow})
            if (x && std::find(t.begin(), t.end(), x) == t.end()) t.push_back(x);
    return t;
}
static bool shares_target(const std::vector<hana_target*>& a, const std::vector<hana_target*>& b) {
    for (hana_target* x : a)
        if (std::find(b.begin(), b.end(), x) != b.end()) return true;
    return false;
}
// Notes work on slot k's stream, complete at slot_done[k], that touches
// targets t, for every other slot to wait for before it touches them.
static void note_pending(hana_ctx* c, int k, const std::vector<hana_target*>& t) {
    for (int b = 0; b < c->nslots; b++)
        if (b != k)
            for (hana_target* x : t)
                if (std::find(c->pend[k][b].begin(), c->pend[k][b].end(), x) == c->pend[k][b].end())
                    c->pend[k][b].push_back(x);
}

int ensure(hana_ctx* c, DevBuf& b, size_t bytes) {
    if (bytes <= b.cap) return HANA_OK;
    if (b.p) {
        int r = sync_all(c);
        if (r) return r;
        HIP_TRY(c, hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
    }
    size_t nb = std::max(bytes, (size_t)256);
    hipError_t e = hipMalloc(&b.p, nb);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, HANA_E_NOMEM, "device allocation of " + std::to_string(nb) + " bytes failed");
    }
    b.cap = nb;
    return HANA_OK;
}

uint32_t pack_bytes(const uint8_t rgba[4]) {
    return (uint32_t)rgba[0] | ((uint32_t)rgba[1] << 8) | ((uint32_t)rgba[2] << 16) | ((uint32_t)rgba[3] << 24);
}

int materialize_clear(hana_ctx* c, hana_target* t, hipStream_t s) {
    if (!t->pending_clear) return HANA_OK;
    size_t n = (size_t)t->w * t->h;
    HIP_TRY(c, launch_fill32(t->color, t->clear_rgba, n, s));
    uint32_t db;
    std::memcpy(&db, &t->clear_depth, 4);
    HIP_TRY(c, launch_fill32(reinterpret_cast<uint32_t*>(t->depth), db, n, s));
    t->pending_clear = false;
    return HANA_OK;
}

// Lays out the workspace of a set of passes.  Passes are ordered (stably) by
// raster group, and within the main group those that sample a shadow map
// drawn by the first group come last: the persistent raster's queue holds the
// tiles in pass order, so every shadow tile precedes the tiles that wait for
// them.  Every per-pass slice is 1024-vertex aligned.
int build_plan(hana_ctx* c, std::vector<PassSpec> specs, Plan& pl, const Tuning& tu) {
    {
        std::vector<std::pair<int, size_t>> key(specs.size());
        for (size_t i = 0; i < specs.size(); i++) {
            const PassSpec& s = specs[i];
            bool dep = false;
            if (s.shadow && s.group == 1)
                for (const PassSpec& o : specs) dep = dep || (o.group == 0 && o.target == s.shadow);
            key[i] = {s.group * 2 + (dep ? 1 : 0), i};
        }
        std::stable_sort(key.begin(), key.end(),
                         [](const std::pair<int, size_t>& a, const std::pair<int, size_t>& b) { return a.first < b.first; });
        std::vector<PassSpec> sorted;
        sorted.reserve(specs.size());
        for (auto& k : key) sorted.push_back(specs[k.second]);
        pl = Plan();
        specs.swap(sorted);
        // the first pass that waits for the shadow tiles
        size_t first_dep = specs.size();
        for (size_t i = 0; i < key.size(); i++)
            if (key[i].first == 3) { first_dep = i; break; }
        pl.q_dep_from = (uint32_t)first_dep;   // a pass index until the tiles are known (below)
    }
    const int tshift = tu.tile_shift, tedge = tile_edge(tu.tile_shift);
    pl.specs = specs;
    int n = (int)specs.size();
    pl.descs.resize(n);
    pl.unis.resize(n);
    size_t voff = 0, vsoff = 0, fo = 0;
    int vb = 0, fb = 0, tile0 = 0;
    int tb[2] = {0, 0};
    pl.small_tris = true;
    for (int i = 0; i < n; i++) {
        const PassSpec& s = specs[i];
        if (s.group < 0 || s.group > 1) return fail(c, HANA_E_INVALID, "bad pass group");
        PassDesc& d = pl.descs[i];
        std::memset(&d, 0, sizeof d);
        derive_dev_uniforms(s.u, pl.unis[i]);
        hana_mesh* m = s.mesh;
        d.nverts = m->V;
        d.nfaces = m->F;
        d.kind = s.kind;
        d.nvary = (s.kind == SHADER_GROUND || s.kind == SHADER_TOON || s.kind == SHADER_SHADOW) ? 1
                  : (s.kind == SHADER_TEXTURE ? 2 : 8);
        size_t V = (size_t)m->V;
        d.px = m->soa; d.py = m->soa + V; d.pz = m->soa + 2 * V;
        d.nx = m->soa + 3 * V; d.ny = m->soa + 4 * V; d.nz = m->soa + 5 * V;
        d.tu = m->soa + 6 * V; d.tv = m->soa + 7 * V;
        d.idx = m->idx;
        if (s.diffuse) { d.diffuse = s.diffuse->texels; d.dw = s.diffuse->w; d.dh = s.diffuse->h; }
        if (s.normal) { d.normal = s.normal->texels; d.nw = s.normal->w; d.nh = s.normal->h; }
        if (s.shadow) {
            d.shadow = s.shadow->color;
            d.shadow_depth = s.shadow->depth;
            d.sw = s.shadow->w;
            d.sh = s.shadow->h;
        }
        hana_target* t = s.target;
        d.color = t->color; d.depth = t->depth; d.winner = t->winner;
        d.width = t->w; d.height = t->h;
        d.tiles_x = (t->w + tedge - 1) / tedge;
        d.tiles_y = (t->h + tedge - 1) / tedge;
        d.share_small = (int64_t)t->w * t->h > 4 * (int64_t)m->F;
        if (d.share_small) pl.small_tris = false;
        const size_t vchunk = (size_t)BLOCK * VPT, fchunk = (size_t)BLOCK * FPT;
        size_t vpad = (V + vchunk - 1) / vchunk * vchunk;
        // slices: pointers are resolved by bind_plan once the arena exists
        pl.v_off.push_back(voff);
        pl.vs_off.push_back(vsoff);
        pl.f_off.push_back(fo);
        voff += vpad;
        vsoff += vpad * (size_t)d.nvary;
        fo += (size_t)m->F;
        d.vblock0 = vb;
        vb += (int)(vpad / vchunk);
        d.fblock0 = fb;
        fb += (int)(((size_t)m->F + fchunk - 1) / fchunk);
        d.tile0 = tile0;
        tile0 += d.tiles_x * d.tiles_y;
        d.tblock0 = tb[s.group];
        tb[s.group] += d.tiles_x * d.tiles_y;
    }
    pl.vblocks = vb;
    pl.fblocks = fb;
    pl.total_v = voff;
    pl.total_vs = vsoff;
    pl.total_f = fo;
    pl.total_tiles = tile0;
    pl.tile_shift = tshift;
    pl.raster_waves = tu.raster_waves;
    pl.raster_wgs = tu.raster_wgs_per_cu > 0 ? tu.raster_wgs_per_cu * c->num_cus : 0;
    pl.q_g0 = (uint32_t)tb[0];
    pl.q_dep_from = pl.q_dep_from < (uint32_t)n ? (uint32_t)pl.descs[pl.q_dep_from].tile0 : (uint32_t)tile0;
    pl.ngroups = 0;
    for (int g = 0; g < 2; g++) {
        int first = -1, count = 0;
        for (int i = 0; i < n; i++)
            if (specs[i].group == g) {
                if (first < 0) first = i;
                count++;
            }
        pl.group_first[g] = first < 0 ? 0 : first;
        pl.group_count[g] = count;
        pl.group_blocks[g] = tb[g];
    }
    // k_raster's pass of each run of 2^TILE_PASS_SHIFT workgroups (a group's
    // passes are contiguous, in tblock0 order): one scalar load per workgroup
    // instead of a binary search over the group's descriptors
    pl.tile_pass.clear();
    for (int g = 0; g < 2; g++) {
        pl.tp_first[g] = (int)pl.tile_pass.size();
        const int first = pl.group_first[g], cnt = pl.group_count[g], nb = pl.group_blocks[g];
        const int run = 1 << TILE_PASS_SHIFT;
        int p = 0;
        for (int b0 = 0; cnt > 0 && b0 < nb; b0 += run) {
            while (p + 1 < cnt && pl.descs[first + p + 1].tblock0 <= b0) p++;
            const int last = std::min(b0 + run, nb) - 1;
            const bool pure = !(p + 1 < cnt && pl.descs[first + p + 1].tblock0 <= last);
            const PassDesc& d = pl.descs[first + p];
            pl.tile_pass.push_back(make_uint2((uint32_t)p | (pure ? TILE_PASS_PURE : 0u), (uint32_t)(d.tile0 - d.tblock0)));
        }
    }
    return HANA_OK;
}

// Allocates the arena for a plan and resolves the slice pointers.
int bind_plan(hana_ctx* c, Plan& pl, Arena& A, uint32_t min_bins, uint32_t min_polys, uint32_t min_tris,
              bool with_stats, hipStream_t s) {
    int n = (int)pl.descs.size();
    int r;
    pl.uni_off = (sizeof(PassDesc) * std::max(n, 1) + 255) & ~(size_t)255;
    pl.tp_off = (pl.uni_off + sizeof(DevUniforms) * std::max(n, 1) + 255) & ~(size_t)255;
    if ((r = ensure(c, A.tables, pl.tp_off + sizeof(uint2) * std::max<size_t>(pl.tile_pass.size(), 1)))) return r;
    if ((r = ensure(c, A.vr, sizeof(float4) * std::max<size_t>(pl.total_v, 1)))) return r;
    if ((r = ensure(c, A.vs, sizeof(float) * std::max<size_t>(pl.total_vs, 1)))) return r;
    if ((r = ensure(c, A.code, std::max<size_t>(pl.total_v, 4)))) return r;
    if ((r = ensure(c, A.pof, sizeof(int32_t) * std::max<size_t>(pl.total_f, 1)))) return r;
    size_t nt = (size_t)pl.total_tiles + 1;
    if ((r = ensure(c, A.tile_count, sizeof(uint32_t) * nt))) return r;
    if ((r = ensure(c, A.tile_offset, sizeof(uint32_t) * nt))) return r;
    if ((r = ensure(c, A.tile_cursor, sizeof(uint32_t) * nt))) return r;
    if ((r = ensure(c, A.scan_part, sizeof(uint64_t) * (size_t)scan_chunks(pl.total_tiles)))) return r;
    // (triangle, tile) entries: most triangles touch one tile and about half
    // are culled; the scan reports the exact need if this is exceeded
    uint64_t want_bins = std::max<uint64_t>({(uint64_t)min_bins, (uint64_t)4 * pl.total_f + 2 * nt, 1u << 22});
    want_bins = std::min<uint64_t>(want_bins, MAX_PAIRS);
    if (want_bins > A.bin_cap) {
        if ((r = ensure(c, A.bins, sizeof(uint32_t) * want_bins))) return r;
        A.bin_cap = (uint32_t)want_bins;
    }
    uint64_t want_polys = std::max<uint64_t>({(uint64_t)min_polys, pl.total_f / 16 + 4096});
    want_polys = std::min<uint64_t>(want_polys, POOL_CAP_MAX);
    if (want_polys > A.poly_cap) {
        if ((r = ensure(c, A.pool_vr, sizeof(float4) * POLY_MAX * want_polys))) return r;
        if ((r = ensure(c, A.pool_vs, sizeof(float) * POLY_MAX * POLY_VARY * want_polys))) return r;
        if ((r = ensure(c, A.pool_face, sizeof(int32_t) * want_polys))) return r;
        if ((r = ensure(c, A.pool_pass, sizeof(int32_t) * want_polys))) return r;
        if ((r = ensure(c, A.pool_n, sizeof(int32_t) * want_polys))) return r;
        A.poly_cap = (uint32_t)want_polys;
    }
    // listed triangles: a segment of BLOCK*FPT records per k_setup workgroup,
    // then at most eight fan triangles per clipped polygon
    uint64_t seg_region = (uint64_t)pl.fblocks * BLOCK * FPT;
    uint64_t want_tris = seg_region + std::max<uint64_t>({(uint64_t)min_tris, 8ull * A.poly_cap, 1024});
    if (want_tris > 0xFFFFFFF0ull) return fail(c, HANA_E_NOMEM, "launch sequence too large (triangle list)");
    if (want_tris > A.tris_cap) {
        if ((r = ensure(c, A.tris, sizeof(uint4) * want_tris))) return r;
        A.tris_cap = (uint32_t)want_tris;
    }
    if ((r = ensure(c, A.seg_count, sizeof(uint32_t) * std::max(pl.fblocks, 1)))) return r;
    // k_setup's per-segment tile lists, as long as its LDS table (launch_setup)
    uint32_t seg_stride = pl.small_tris ? 512u : 1024u;
    size_t nsegs = (size_t)std::max(pl.fblocks, 1);
    if ((r = ensure(c, A.seg_tiles, sizeof(uint2) * seg_stride * nsegs))) return r;
    if ((r = ensure(c, A.seg_wide, sizeof(uint16_t) * BLOCK * FPT * nsegs))) return r;
    if ((r = ensure(c, A.seg_info, sizeof(uint32_t) * nsegs))) return r;
    if ((r = ensure(c, A.qstate, 256))) return r;
    // every face may fail the trivial accept (camera inside the mesh)
    uint64_t want_q = std::max<uint64_t>(pl.total_f, 1024);
    want_q = std::min<uint64_t>(want_q, 0xFFFFFFF0ull);
    if (want_q > A.clipq_cap) {
        if ((r = ensure(c, A.clip_queue, sizeof(uint64_t) * want_q))) return r;
        A.clipq_cap = (uint32_t)want_q;
    }
    bool fresh_status = A.status.p == nullptr;
    if ((r = ensure(c, A.status, sizeof(uint32_t) * STATUS_WORDS))) return r;
    if (fresh_status) HIP_TRY(c, launch_fill32(A.status.as<uint32_t>(), 0u, STATUS_WORDS, s));
    if ((r = ensure(c, A.stats, sizeof(uint64_t) * ST_WORDS * std::max(n, 1)))) return r;
    for (int i = 0; i < n; i++) {
        PassDesc& d = pl.descs[i];
        size_t voff = pl.v_off[i], vsoff = pl.vs_off[i], fo = pl.f_off[i];
        d.vr = A.vr.as<float4>() + voff;
        d.vs = A.vs.as<float>() + vsoff;
        d.code = A.code.as<uint8_t>() + voff;
        d.poly_of_face = A.pof.as<int32_t>() + fo;
        d.u = reinterpret_cast<const DevUniforms*>(A.tables.as<char>() + pl.uni_off) + i;
        d.stats = with_stats ? A.stats.as<uint64_t>() + (size_t)ST_WORDS * i : nullptr;
    }
    Workspace& ws = pl.ws;
    ws.tile_count = A.tile_count.as<uint32_t>();
    ws.tile_offset = A.tile_offset.as<uint32_t>();
    ws.tile_cursor = A.tile_cursor.as<uint32_t>();
    ws.bins = A.bins.as<uint32_t>();
    ws.pool_vr = A.pool_vr.as<float4>();
    ws.pool_vs = A.pool_vs.as<float>();
    ws.pool_face = A.pool_face.as<int32_t>();
    ws.pool_pass = A.pool_pass.as<int32_t>();
    ws.pool_n = A.pool_n.as<int32_t>();
    ws.clip_queue = A.clip_queue.as<uint64_t>();
    ws.tris = A.tris.as<uint4>();
    ws.tris_cap = A.tris_cap;
    ws.seg_region = (uint32_t)seg_region;
    ws.seg_count = A.seg_count.as<uint32_t>();
    ws.seg_tiles = A.seg_tiles.as<uint2>();
    ws.seg_wide = A.seg_wide.as<uint16_t>();
    ws.seg_info = A.seg_info.as<uint32_t>();
    ws.seg_stride = seg_stride;
    ws.scan_part = A.scan_part.as<uint64_t>();
    ws.clipq_cap = A.clipq_cap;
    ws.status = A.status.as<uint32_t>();
    ws.status_out = nullptr;
    ws.ntiles = pl.total_tiles;
    ws.bin_cap = A.bin_cap;
    ws.poly_cap = A.poly_cap;
    ws.qstate = A.qstate.as<uint32_t>();
    ws.q_items = (uint32_t)pl.total_tiles;
    ws.q_g0 = pl.q_g0;
    ws.q_dep_from = pl.q_dep_from;
    ws.tile_shift = pl.tile_shift;
    ws.raster_waves = pl.raster_waves;
    for (int g = 0; g < 2; g++)
        ws.tile_pass[g] = pl.group_count[g] > 0
                              ? reinterpret_cast<const uint2*>(A.tables.as<char>() + pl.tp_off) + pl.tp_first[g]
                              : nullptr;
    return HANA_OK;
}

// Resolves clear flags of the plan's targets (fused into their first pass).
void take_clears(Plan& pl) {
    for (size_t i = 0; i < pl.descs.size(); i++) {
        hana_target* t = pl.specs[i].target;
        PassDesc& d = pl.descs[i];
        d.clear = 0;
        if (t->pending_clear) {
            d.clear = 1;
            d.clear_rgba = t->clear_rgba;
            d.clear_depth = t->clear_depth;
            t->pending_clear = false;
        }
    }
}

hipEvent_t pool_event(hana_ctx* c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Records an event pair around one launch when timing is on (never inside a
// graph capture: `timed` is false there).
struct StageTimer {
    hana_ctx* c;
    int stage;
    bool on;
    hipEvent_t a = nullptr, b = nullptr;
    StageTimer(hana_ctx* c_, int st, bool timed, hipStream_t s) : c(c_), stage(st), on(timed && c_->timing) {
        if (on) {
            a = pool_event(c);
            b = pool_event(c);
            on = a && b && hipEventRecord(a, s) == hipSuccess;
        }
    }
    void done(hipStream_t s) {
        if (on && hipEventRecord(b, s) == hipSuccess) c->ev_pending.push_back({stage, {a, b}});
    }
};

// The geometry stages: vertex, set-up, clipping, tile scan, binning.
int enqueue_geometry(hana_ctx* c, const Plan& pl, const Arena& A, hipStream_t s, bool timed) {
    const PassDesc* d = A.tables.as<PassDesc>();
    int n = (int)pl.descs.size();
    {
        StageTimer t(c, 0, timed, s);
        HIP_TRY(c, launch_vertex(d, n, pl.ws, pl.vblocks, s));
        t.done(s);
    }
    {
        StageTimer t(c, 1, timed, s);
        HIP_TRY(c, launch_setup(d, n, pl.ws, pl.fblocks, pl.small_tris, s));
        t.done(s);
    }
    {
        StageTimer t(c, 2, timed, s);
        HIP_TRY(c, launch_clip(d, pl.ws, s));
        t.done(s);
    }
    {
        StageTimer t(c, 3, timed, s);
        HIP_TRY(c, launch_scan(pl.ws, s));
        t.done(s);
    }
    {
        StageTimer t(c, 4, timed, s);
        HIP_TRY(c, launch_bin(d, pl.ws, pl.fblocks, pl.small_tris, s));
        t.done(s);
    }
    return HANA_OK;
}

// The raster stages: one launch per dependency group, or one persistent
// launch for both (timed as the first group's stage).
int enqueue_raster(hana_ctx* c, const Plan& pl, const Arena& A, hipStream_t s, bool timed) {
    const PassDesc* d = A.tables.as<PassDesc>();
    if (pl.raster_wgs > 0) {
        StageTimer t(c, 6, timed, s);
        HIP_TRY(c, launch_raster_queue(d, (int)pl.descs.size(), pl.ws, pl.raster_wgs, s));
        t.done(s);
        return HANA_OK;
    }
    for (int g = 0; g < 2; g++)
        if (pl.group_count[g]) {
            StageTimer t(c, 6 + g, timed, s);
            HIP_TRY(c, launch_raster(d + pl.group_first[g], pl.group_count[g], pl.ws, pl.group_blocks[g], g, s));
            t.done(s);
        }
    return HANA_OK;
}

int enqueue_sequence(hana_ctx* c, const Plan& pl, const Arena& A, hipStream_t s, bool timed = true) {
    // every kernel reads the workspace: a plan that was never bound to an
    // arena must not reach the device
    if (pl.descs.empty() || !pl.ws.status || !A.tables.p)
        return fail(c, HANA_E_INVALID, "internal error: a sequence without a bound plan");
    int r = enqueue_geometry(c, pl, A, s, timed);
    return r ? r : enqueue_raster(c, pl, A, s, timed);
}

// Uploads descriptors and uniforms through a pinned staging slot.
// Stages the plan's descriptors and uniforms in the pinned ring and copies
// them into the arena's tables buffer (one copy).
int upload_tables(hana_ctx* c, const Plan& pl, Arena& A, hipStream_t s) {
    size_t nd = sizeof(PassDesc) * pl.descs.size(), nu = sizeof(DevUniforms) * pl.unis.size();
    size_t nt = sizeof(uint2) * pl.tile_pass.size();
    size_t need = pl.tp_off + nt;
    if (need > c->ring_bytes) {
        int r = sync_all(c);
        if (r) return r;
        for (int i = 0; i < kRing; i++) {
            if (c->ring_host[i]) HIP_TRY(c, hipHostFree(c->ring_host[i]));
            c->ring_host[i] = nullptr;
            c->ring_used[i] = false;
        }
        c->ring_bytes = 0;
        for (int i = 0; i < kRing; i++) HIP_TRY(c, hipHostMalloc(&c->ring_host[i], need * 2));
        c->ring_bytes = need * 2;
    }
    int slot = c->ring_next;
    c->ring_next = (c->ring_next + 1) % kRing;
    if (c->ring_used[slot]) HIP_TRY(c, hipEventSynchronize(c->ring_event[slot]));
    char* h = static_cast<char*>(c->ring_host[slot]);
    std::memcpy(h, pl.descs.data(), nd);
    std::memcpy(h + pl.uni_off, pl.unis.data(), nu);
    if (nt) std::memcpy(h + pl.tp_off, pl.tile_pass.data(), nt);
    HIP_TRY(c, hipMemcpyAsync(A.tables.p, h, need, hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipEventRecord(c->ring_event[slot], s));
    c->ring_used[slot] = true;
    return HANA_OK;
}

// Grows an arena's bins, clip pool and triangle list to the needs reported
// by an overflow (capacities never shrink).
static int grow_arena(hana_ctx* c, Arena& A, uint32_t flags, uint32_t need_bins, uint32_t need_polys,
                      uint32_t need_tris) {
    if (flags & FLAG_BIN_OVERFLOW) {
        uint64_t want = std::max<uint64_t>((uint64_t)need_bins + need_bins / 4 + 1024, A.bin_cap);
        want = std::min<uint64_t>(want, MAX_PAIRS);
        int r = ensure(c, A.bins, sizeof(uint32_t) * want);
        if (r) return r;
        A.bin_cap = (uint32_t)want;
    }
    if (flags & FLAG_TRIS_OVERFLOW) {
        uint64_t want = std::min<uint64_t>((uint64_t)A.tris_cap + need_tris + need_tris / 4 + 1024, 0xFFFFFFF0ull);
        int r = ensure(c, A.tris, sizeof(uint4) * want);
        if (r) return r;
        A.tris_cap = (uint32_t)want;
    }
    if (flags & FLAG_POLY_OVERFLOW) {
        uint64_t want = std::max<uint64_t>((uint64_t)need_polys + need_polys / 4 + 1024, A.poly_cap);
        want = std::min<uint64_t>(want, POOL_CAP_MAX);
        int r;
        if ((r = ensure(c, A.pool_vr, sizeof(float4) * POLY_MAX * want))) return r;
        if ((r = ensure(c, A.pool_vs, sizeof(float) * POLY_MAX * POLY_VARY * want))) return r;
        if ((r = ensure(c, A.pool_face, sizeof(int32_t) * want))) return r;
        if ((r = ensure(c, A.pool_pass, sizeof(int32_t) * want))) return r;
        if ((r = ensure(c, A.pool_n, sizeof(int32_t) * want))) return r;
        A.poly_cap = (uint32_t)want;
    }
    return HANA_OK;
}

// Renders a logged call again on the context stream with slot 0's workspace
// (every stream is idle: the status check has waited for them).
static int replay_call(hana_ctx* c, LoggedCall& L) {
    int r;
    Plan& pl = L.pl;
    Arena& A = slot_arena(c, 0);
    if ((r = bind_plan(c, pl, A, 0, 0, 0, false, c->stream))) return r;
    pl.ws.status_out = c->status_dev[0];
    if ((r = upload_tables(c, pl, A, c->stream))) return r;
    if ((r = enqueue_geometry(c, pl, A, c->stream, false))) return r;
    for (const Fill& f : L.fills) {
        size_t n = (size_t)f.target->w * f.target->h;
        HIP_TRY(c, launch_fill32(f.target->color, f.rgba, n, c->stream));
        uint32_t db;
        std::memcpy(&db, &f.depth, 4);
        HIP_TRY(c, launch_fill32(reinterpret_cast<uint32_t*>(f.target->depth), db, n, c->stream));
    }
    return enqueue_raster(c, pl, A, c->stream, false);
}

// Reads the sticky status of every sequence since the last check, over all
// draw slots (waits for every stream).
// - A capacity overflow (bins, clip pool, clip queue, triangle list) left its
//   sequence's targets untouched.  Every slot's workspace is grown to the
//   largest need seen and the calls since the last check are rendered again
//   in call order; a draw over the same state twice leaves it as drawn once
//   (the depth test keeps the last of the fragments at the minimum depth), so
//   the targets end as the calls in order leave them.  With replay_overflow
//   off, the caller is asked to render again (HANA_E_OVERFLOW).
// - More (triangle, tile) pairs than 32-bit bins reach: that sequence was not
//   rasterised (HANA_E_NOMEM).
// - A clipped polygon with more than POLY_MAX vertices was dropped, the rest
//   drawn (HANA_E_UNSUPPORTED).
int check_status(hana_ctx* c) {
    if (!c->status_pending && !c->noalloc_pending) return HANA_OK;
    int r;
    uint32_t seen = 0;
    for (int attempt = 0;; attempt++) {
        if ((r = sync_all(c))) return r;
        uint32_t flags = 0, need_bins = 0, need_polys = 0, need_tris = 0;
        for (int k = 0; k < c->nslots; k++) {
            const uint32_t* host = slot_status(c, k);
            if (!host) continue;
            flags |= host[STATUS_FLAGS];
            need_bins = std::max(need_bins, host[STATUS_NEED_BINS]);
            need_polys = std::max(need_polys, host[STATUS_NEED_POLYS]);
            need_tris = std::max(need_tris, host[STATUS_NEED_TRIS]);
        }
        seen |= flags;
        if (flags) {
            for (int k = 0; k < c->nslots; k++) {
                uint32_t* host = slot_status(c, k);
                if (!host) continue;
                Arena& A = slot_arena(c, k);
                if (A.status.p)
                    HIP_TRY(c, launch_fill32(A.status.as<uint32_t>() + STATUS_SEQ_WORDS, 0u,
                                             STATUS_WORDS - STATUS_SEQ_WORDS, c->stream));
                std::memset(host + STATUS_SEQ_WORDS, 0, sizeof(uint32_t) * (STATUS_WORDS - STATUS_SEQ_WORDS));
            }
            // the next geometry stages may run on any stream
            HIP_TRY(c, hipStreamSynchronize(c->stream));
        }
        uint32_t cap = flags & FLAG_CAPACITY;
        if (cap)
            for (int k = 0; k < c->nslots; k++)
                if ((r = grow_arena(c, slot_arena(c, k), flags, need_bins, need_polys, need_tris))) {
                    c->log.clear();
                    c->status_pending = false;
                    return r;
                }
        if (cap && c->replay_overflow && attempt < 3) {
            c->replays++;
            for (LoggedCall& L : c->log)
                if ((r = replay_call(c, L))) {
                    c->log.clear();
                    c->status_pending = false;
                    return r;
                }
            continue;   // wait for the replay and check it
        }
        c->log.clear();
        c->status_pending = false;
        if (cap)
            return fail(c, HANA_E_OVERFLOW,
                        c->replay_overflow ? "workspace capacity still exceeded after growing it three times"
                                           : "bin or clip-pool capacity exceeded since the last sync; buffers were "
                                             "grown, render those frames again");
        break;
    }
    // no-alloc frames: no workspace to grow, their errors are reported
    uint32_t na = 0;
    if (c->noalloc_pending) {
        na = c->noalloc_status_host[STATUS_FLAGS];
        std::memset(c->noalloc_status_host, 0, sizeof(uint32_t) * STATUS_WORDS);
        c->noalloc_pending = false;
    }
    if ((seen | na) & FLAG_QUEUE_STALL)
        return fail(c, HANA_E_DEVICE, "the persistent raster's wait for the shadow tiles timed out");
    if (na & FLAG_CAPACITY)
        return fail(c, HANA_E_OVERFLOW,
                    "a hana_render_frame_noalloc frame outgrew the bins or clip pool of the no-alloc scratch and left "
                    "its targets untouched; render it with hana_render_frame");
    seen |= na;
    if (seen & FLAG_PAIR_LIMIT)
        return fail(c, HANA_E_NOMEM,
                    "a draw, frame or batch part covered more than 4,294,967,280 (triangle, tile) pairs and was not "
                    "rasterised; split it");
    if (seen & FLAG_POLY_TOO_LARGE)
        return fail(c, HANA_E_UNSUPPORTED,
                    "a clipped polygon had more than " + std::to_string(POLY_MAX) +
                        " vertices and was dropped; the rest was drawn");
    return HANA_OK;
}

// Renders a list of passes (per-call path).  Consecutive calls rotate over
// the context's slots, each a stream with its own workspace: the geometry
// stages of a call run while earlier calls rasterise, and its raster waits
// for the previous call to finish, so targets and shadow maps see the calls
// in order.  With timing or statistics on, every call uses slot 0.
int run_passes(hana_ctx* c, std::vector<PassSpec> specs) {
    int r;
    if (c->log.size() >= kMaxLoggedCalls && (r = check_status(c))) return r;
    const std::vector<hana_target*> touched = touched_targets(specs);
    std::vector<hana_target*> sampled;   // sampled as a shadow map, not drawn here
    for (auto& s : specs) {
        if (!s.shadow || !s.shadow->pending_clear) continue;
        bool drawn = false;
        for (auto& o : specs) drawn = drawn || o.target == s.shadow;
        if (!drawn && std::find(sampled.begin(), sampled.end(), s.shadow) == sampled.end()) sampled.push_back(s.shadow);
    }
    Plan pl;
    if ((r = build_plan(c, std::move(specs), pl, c->tuning))) return r;
    bool rotate = c->pipelined && !c->timing && !c->stats_enabled;
    // Large sequences rotate over two slots: with a third in flight, the 2.6 M
    // faces of a 1.3 M-triangle frame with its shadow pass ran 1.2 % slower,
    // while frames of up to 400 k faces ran 6 % faster.
    int nslots = pl.total_f >= kLargeSequenceFaces ? std::min(2, c->nslots) : c->nslots;
    int a = rotate ? c->next_slot % nslots : 0;
    hipStream_t S = slot_stream(c, a);
    Arena& A = slot_arena(c, a);
    // the slots' workspaces grow together, so that a sequence of calls
    // allocates (and waits for the streams) once, not once per slot
    if (rotate)
        for (int k = 0; k < nslots; k++)
            if (k != a && (r = bind_plan(c, pl, slot_arena(c, k), 0, 0, 0, false, slot_stream(c, k)))) return r;
    if ((r = bind_plan(c, pl, A, 0, 0, 0, c->stats_enabled, S))) return r;
    pl.ws.status_out = c->status_dev[a];   // the scan leaves the status words in host memory
    // a target sampled as a shadow map by a pass of the same plan must be drawn
    // by an earlier group; clears fuse into the first pass writing the target
    take_clears(pl);
    if ((r = upload_tables(c, pl, A, S))) return r;
    int n = (int)pl.descs.size();
    if (c->stats_enabled) {
        HIP_TRY(c, launch_fill32(reinterpret_cast<uint32_t*>(A.stats.p), 0u, (size_t)ST_WORDS * 2 * n, S));
    }
    if ((r = enqueue_geometry(c, pl, A, S, true))) return r;
    // the raster waits for the calls in flight on other slots that touch one
    // of its targets; calls on disjoint targets run concurrently.  After a
    // no-alloc frame (context stream, targets not noted) every slot waits once.
    if (a != 0 && (c->noalloc_wait_mask & (1u << a))) {
        HIP_TRY(c, hipStreamWaitEvent(S, c->slot_done[0], 0));
        c->noalloc_wait_mask &= ~(1u << a);
    }
    for (int k = 0; k < c->nslots; k++)
        if (k != a && (c->pend[k][a].size() > kMaxPendingTargets || shares_target(c->pend[k][a], touched))) {
            HIP_TRY(c, hipStreamWaitEvent(S, c->slot_done[k], 0));
            c->pend[k][a].clear();
        }
    std::vector<Fill> fills;
    for (hana_target* t : sampled) {
        fills.push_back({t, t->clear_rgba, t->clear_depth});
        if ((r = materialize_clear(c, t, S))) return r;
    }
    if ((r = enqueue_raster(c, pl, A, S, true))) return r;
    HIP_TRY(c, hipEventRecord(c->slot_done[a], S));
    if (c->stats_enabled) {
        if (c->stats_passes < n) {
            HIP_TRY(c, hipStreamSynchronize(S));   // the previous call may still copy into it
            if (c->stats_host) HIP_TRY(c, hipHostFree(c->stats_host));
            c->stats_host = nullptr;
            c->stats_passes = 0;
            HIP_TRY(c, hipHostMalloc((void**)&c->stats_host, sizeof(uint64_t) * ST_WORDS * n));
        }
        HIP_TRY(c, hipMemcpyAsync(c->stats_host, A.stats.p, sizeof(uint64_t) * ST_WORDS * n,
                                  hipMemcpyDeviceToHost, S));
        c->stats_passes = n;
    }
    if (c->pipelined) note_pending(c, a, touched);   // else no other slot ever draws
    c->next_slot = (a + 1) % nslots;
    c->status_pending = true;
    c->log.push_back({std::move(pl), std::move(fills)});
    return HANA_OK;
}

bool valid_kind(int k) { return k >= HANA_SHADER_GROUND && k <= HANA_SHADER_SHADOW; }

int tuning_from(hana_ctx* c, const hana_tuning* t, Tuning& out) {
    if (!t || !(t->tile_size == 16 || t->tile_size == 32) || t->raster_wgs_per_cu < 0 || t->raster_wgs_per_cu > 16 ||
        t->parts < 1 || t->parts > (int)kMaxBatchParts || !(t->raster_waves == 0 || t->raster_waves == 2 || t->raster_waves == 4))
        return fail(c, HANA_E_INVALID, "tuning: tile_size 16 or 32, raster_wgs_per_cu 0-16, parts 1-4, raster_waves 0, 2 or 4");
    out.tile_shift = t->tile_size == 32 ? 5 : 4;
    out.raster_wgs_per_cu = t->raster_wgs_per_cu;
    out.raster_waves = t->raster_waves ? t->raster_waves : 4;
    out.parts = t->parts;
    return HANA_OK;
}

}  // namespace rt
}  // namespace hana

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int hana_gpu_create(int device, hana_ctx** out) { return create_context(device, kDrawSlots, out); }

}  // extern "C"

namespace hana {
namespace rt {

int create_context(int device, int nslots, hana_ctx** out) {
    if (!out || nslots < 1 || nslots > kDrawSlots) return HANA_E_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return HANA_E_DEVICE;
    if (device < 0 || device >= count) return HANA_E_INVALID;
    hana_ctx* c = new hana_ctx();
    c->device = device;
    c->nslots = nslots;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        c->num_cus = cus;
    auto undo = [c]() {
        if (c->noalloc_status_host) (void)hipHostFree(c->noalloc_status_host);
        for (int i = 0; i < kRing; i++)
            if (c->ring_event[i]) (void)hipEventDestroy(c->ring_event[i]);
        for (int k = 0; k < c->nslots; k++) {
            if (c->slot_done[k]) (void)hipEventDestroy(c->slot_done[k]);
            if (slot_status(c, k)) (void)hipHostFree(slot_status(c, k));
            if (slot_stream(c, k)) (void)hipStreamDestroy(slot_stream(c, k));
        }
        (void)hipGetLastError();
        delete c;
        return HANA_E_DEVICE;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return undo();
    for (int i = 0; i < kRing; i++)
        if (hipEventCreateWithFlags(&c->ring_event[i], hipEventDisableTiming) != hipSuccess) return undo();
    for (int k = 0; k < c->nslots; k++) {
        if (k && hipStreamCreateWithFlags(&c->xstream[k - 1], hipStreamNonBlocking) != hipSuccess) return undo();
        uint32_t*& host = k ? c->xstatus_host[k - 1] : c->status_host;
        if (hipHostMalloc((void**)&host, sizeof(uint32_t) * STATUS_WORDS) != hipSuccess) return undo();
        std::memset(host, 0, sizeof(uint32_t) * STATUS_WORDS);
        if (hipHostGetDevicePointer((void**)&c->status_dev[k], host, 0) != hipSuccess) return undo();
        if (hipEventCreateWithFlags(&c->slot_done[k], hipEventDisableTiming) != hipSuccess) return undo();
    }
    if (hipHostMalloc((void**)&c->noalloc_status_host, sizeof(uint32_t) * STATUS_WORDS) != hipSuccess) return undo();
    std::memset(c->noalloc_status_host, 0, sizeof(uint32_t) * STATUS_WORDS);
    if (hipHostGetDevicePointer((void**)&c->noalloc_status_dev, c->noalloc_status_host, 0) != hipSuccess ||
        noalloc_init(c) != HANA_OK)
        return undo();
    *out = c;
    return HANA_OK;
}

}  // namespace rt
}  // namespace hana

extern "C" {

static void free_arena(Arena& a) {
    DevBuf* bufs[] = {&a.tables, &a.vr, &a.vs, &a.code, &a.tris, &a.seg_count, &a.seg_tiles, &a.seg_wide,
                      &a.seg_info, &a.qstate, &a.scan_part, &a.pof, &a.tile_count, &a.tile_offset, &a.tile_cursor,
                      &a.bins, &a.pool_vr, &a.pool_vs, &a.pool_face, &a.pool_pass, &a.pool_n, &a.clip_queue,
                      &a.status, &a.stats};
    for (DevBuf* b : bufs)
        if (b->p) {
            (void)hipFree(b->p);
            b->p = nullptr;
            b->cap = 0;
        }
    a.bin_cap = a.poly_cap = a.clipq_cap = a.tris_cap = 0;
}

int hana_gpu_destroy(hana_ctx* c) {
    if (!c) return HANA_E_INVALID;
    (void)hipSetDevice(c->device);
    for (int k = 0; k < c->nslots; k++) {
        (void)hipStreamSynchronize(slot_stream(c, k));
        free_arena(slot_arena(c, k));
    }
    if (c->present.p) (void)hipFree(c->present.p);
    for (int i = 0; i < kRing; i++) {
        if (c->ring_host[i]) (void)hipHostFree(c->ring_host[i]);
        if (c->ring_event[i]) (void)hipEventDestroy(c->ring_event[i]);
    }
    if (c->stats_host) (void)hipHostFree(c->stats_host);
    if (c->noalloc_status_host) (void)hipHostFree(c->noalloc_status_host);
    for (int k = 0; k < c->nslots; k++) {
        if (slot_status(c, k)) (void)hipHostFree(slot_status(c, k));
        if (c->slot_done[k]) (void)hipEventDestroy(c->slot_done[k]);
    }
    for (auto& p : c->ev_pending) {
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    for (int k = 0; k < c->nslots; k++) (void)hipStreamDestroy(slot_stream(c, k));
    delete c;
    return HANA_OK;
}

const char* hana_gpu_last_error(hana_ctx* c) { return c ? c->err.c_str() : "no context"; }

int hana_gpu_sync(hana_ctx* c) {
    if (!c) return HANA_E_INVALID;
    int r = sync_all(c);
    return r ? r : check_status(c);
}

int hana_gpu_set_timing(hana_ctx* c, int enable) {
    if (!c) return HANA_E_INVALID;
    c->timing = enable != 0;
    return sync_all(c);   // stage times never overlap a draw of the second slot
}

int hana_gpu_read_timing(hana_ctx* c, double* ms, int64_t* launches) {
    if (!c || !ms || !launches) return HANA_E_INVALID;
    int r = sync_all(c);
    if (r) return r;
    for (int i = 0; i < HANA_TIMING_STAGES; i++) {
        ms[i] = 0.0;
        launches[i] = 0;
    }
    for (auto& p : c->ev_pending) {
        float t = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&t, p.second.first, p.second.second));
        ms[p.first] += t;
        launches[p.first] += 1;
        c->ev_pool.push_back(p.second.first);
        c->ev_pool.push_back(p.second.second);
    }
    c->ev_pending.clear();
    return HANA_OK;
}

int hana_gpu_replay_count(hana_ctx* c, int64_t* out) {
    if (!c || !out) return HANA_E_INVALID;
    *out = c->replays;
    return HANA_OK;
}

int hana_gpu_set_stats(hana_ctx* c, int enable) {
    if (!c) return HANA_E_INVALID;
    c->stats_enabled = enable != 0;
    return HANA_OK;
}

int hana_gpu_last_stats(hana_ctx* c, hana_pass_stats* out, int max_passes, int* n_passes) {
    if (!c || !out || !n_passes) return HANA_E_INVALID;
    int r = sync_all(c);
    if (r) return r;
    int n = std::min(max_passes, c->stats_host ? c->stats_passes : 0);
    for (int i = 0; i < n; i++) {
        const uint64_t* s = c->stats_host + (size_t)ST_WORDS * i;
        out[i].faces = (int64_t)s[ST_FACES];
        out[i].clipped_faces = (int64_t)s[ST_CLIPPED];
        out[i].polygons = (int64_t)s[ST_POLYS];
        out[i].tile_entries = (int64_t)s[ST_ENTRIES];
        out[i].pixels_written = (int64_t)s[ST_PIXELS];
    }
    *n_passes = n;
    return HANA_OK;
}

int hana_mesh_create(hana_ctx* c, int32_t nverts, const float* pos, const float* nrm, const float* uv,
                     int32_t nfaces, const int32_t* idx, hana_mesh** out) {
    if (!c || !out || nverts < 0 || nfaces < 0 || (nverts && !pos) || (nfaces && !idx))
        return fail(c, HANA_E_INVALID, "hana_mesh_create: bad arguments");
    if (nfaces > (1 << 27)) return fail(c, HANA_E_INVALID, "hana_mesh_create: at most 2^27 faces");
    for (int64_t i = 0; i < 3 * (int64_t)nfaces; i++)
        if (idx[i] < 0 || idx[i] >= nverts) return fail(c, HANA_E_INVALID, "hana_mesh_create: index out of range");
    size_t V = (size_t)nverts;
    std::vector<float> soa(8 * std::max<size_t>(V, 1), 0.0f);
    for (size_t i = 0; i < V; i++) {
        for (int k = 0; k < 3; k++) soa[k * V + i] = pos[3 * i + k];
        if (nrm)
            for (int k = 0; k < 3; k++) soa[(3 + k) * V + i] = nrm[3 * i + k];
        if (uv)
            for (int k = 0; k < 2; k++) soa[(6 + k) * V + i] = uv[2 * i + k];
    }
    hana_mesh* m = new hana_mesh();
    m->V = nverts;
    m->F = nfaces;
    if (hipMalloc((void**)&m->soa, sizeof(float) * soa.size()) != hipSuccess ||
        hipMalloc((void**)&m->idx, sizeof(int32_t) * 3 * std::max<size_t>(nfaces, 1)) != hipSuccess) {
        (void)hipGetLastError();
        if (m->soa) (void)hipFree(m->soa);
        delete m;
        return fail(c, HANA_E_NOMEM, "hana_mesh_create: device allocation failed");
    }
    hipError_t e = hipMemcpy(m->soa, soa.data(), sizeof(float) * soa.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && nfaces) e = hipMemcpy(m->idx, idx, sizeof(int32_t) * 3 * (size_t)nfaces, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(m->soa);
        (void)hipFree(m->idx);
        delete m;
        return hip_fail(c, e, "hana_mesh_create: upload");
    }
    *out = m;
    return HANA_OK;
}

// Objects a logged call may be rendered again with are freed only after the
// status check (which may replay it); its result is returned.
int hana_mesh_destroy(hana_ctx* c, hana_mesh* m) {
    if (!c || !m) return HANA_E_INVALID;
    int st = check_status(c);
    (void)sync_all(c);
    (void)hipFree(m->soa);
    (void)hipFree(m->idx);
    delete m;
    return st;
}

int hana_texture_create(hana_ctx* c, int32_t w, int32_t h, int32_t bpp, const uint8_t* bytes, hana_texture** out) {
    if (!c || !out || w <= 0 || h <= 0 || !(bpp == 1 || bpp == 3 || bpp == 4) || !bytes)
        return fail(c, HANA_E_INVALID, "hana_texture_create: bad arguments");
    size_t n = (size_t)w * h;
    std::vector<uint32_t> words(n);
    for (size_t i = 0; i < n; i++) {
        uint32_t v = 0;
        for (int k = 0; k < bpp; k++) v |= (uint32_t)bytes[i * bpp + k] << (8 * k);   // TGAColor(p, bpp)
        words[i] = v;
    }
    hana_texture* t = new hana_texture();
    t->w = w;
    t->h = h;
    if (hipMalloc((void**)&t->texels, sizeof(uint32_t) * n) != hipSuccess) {
        (void)hipGetLastError();
        delete t;
        return fail(c, HANA_E_NOMEM, "hana_texture_create: device allocation failed");
    }
    hipError_t e = hipMemcpy(t->texels, words.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(t->texels);
        delete t;
        return hip_fail(c, e, "hana_texture_create: upload");
    }
    *out = t;
    return HANA_OK;
}

int hana_texture_destroy(hana_ctx* c, hana_texture* t) {
    if (!c || !t) return HANA_E_INVALID;
    int st = check_status(c);
    (void)sync_all(c);
    (void)hipFree(t->texels);
    delete t;
    return st;
}

int hana_target_create(hana_ctx* c, int32_t w, int32_t h, hana_target** out) {
    // tile coordinates travel as 16-bit fields: at most 2^16 tiles a side
    if (!c || !out || w <= 0 || h <= 0 || w > (1 << 20) || h > (1 << 20) || (int64_t)w * h > (1ll << 30))
        return fail(c, HANA_E_INVALID, "hana_target_create: bad size");
    hana_target* t = new hana_target();
    t->w = w;
    t->h = h;
    size_t n = (size_t)w * h;
    if (hipMalloc((void**)&t->color, 4 * n) != hipSuccess || hipMalloc((void**)&t->depth, 4 * n) != hipSuccess) {
        (void)hipGetLastError();
        if (t->color) (void)hipFree(t->color);
        delete t;
        return fail(c, HANA_E_NOMEM, "hana_target_create: device allocation failed");
    }
    // RenderBuffer constructor state: colour {0,0,0,1} -> (0,0,0,255), depth 1
    // (SR/renderbuffer.cpp:3-18)
    t->pending_clear = true;
    t->clear_rgba = 0xFF000000u;
    t->clear_depth = 1.0f;
    *out = t;
    return HANA_OK;
}

int hana_target_destroy(hana_ctx* c, hana_target* t) {
    if (!c || !t) return HANA_E_INVALID;
    int st = check_status(c);
    (void)sync_all(c);
    (void)hipFree(t->color);
    (void)hipFree(t->depth);
    if (t->winner) (void)hipFree(t->winner);
    delete t;
    return st;
}

int hana_target_clear(hana_ctx* c, hana_target* t, const uint8_t rgba[4], float depth) {
    if (!c || !t || !rgba) return HANA_E_INVALID;
    t->pending_clear = true;
    t->clear_rgba = pack_bytes(rgba);
    t->clear_depth = depth;
    return HANA_OK;
}

int hana_target_upload(hana_ctx* c, hana_target* t, const uint8_t* rgba, const float* depth) {
    if (!c || !t) return HANA_E_INVALID;
    size_t n = (size_t)t->w * t->h;
    // draws before the upload are complete (replayed if need be) before it lands
    int r = check_status(c);
    if (r) return r;
    r = join(c);
    if (r || (r = materialize_clear(c, t, c->stream))) return r;
    if (rgba) HIP_TRY(c, hipMemcpyAsync(t->color, rgba, 4 * n, hipMemcpyHostToDevice, c->stream));
    if (depth) HIP_TRY(c, hipMemcpyAsync(t->depth, depth, 4 * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return HANA_OK;
}

// The reads first complete every draw (and render again those that outgrew
// the workspace), then copy; a draw's error is returned after the copy.
int hana_target_read(hana_ctx* c, hana_target* t, uint8_t* rgba, float* depth) {
    if (!c || !t) return HANA_E_INVALID;
    int st = check_status(c);
    if (st == HANA_E_DEVICE) return st;
    int r = join(c);
    if (r || (r = materialize_clear(c, t, c->stream))) return r;
    size_t n = (size_t)t->w * t->h;
    if (rgba) HIP_TRY(c, hipMemcpyAsync(rgba, t->color, 4 * n, hipMemcpyDeviceToHost, c->stream));
    if (depth) HIP_TRY(c, hipMemcpyAsync(depth, t->depth, 4 * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return st;
}

int hana_target_present(hana_ctx* c, hana_target* t, uint8_t* bgra) {
    if (!c || !t || !bgra) return fail(c, HANA_E_INVALID, "hana_target_present: null argument");
    int st = check_status(c);
    if (st == HANA_E_DEVICE) return st;
    int r = join(c);
    if (r || (r = materialize_clear(c, t, c->stream))) return r;
    size_t n = (size_t)t->w * t->h;
    if ((r = ensure(c, c->present, 4 * n))) return r;
    uint32_t* dst = c->present.as<uint32_t>();
    HIP_TRY(c, launch_present(t->color, dst, t->w, t->h, c->stream));
    HIP_TRY(c, hipMemcpyAsync(bgra, dst, 4 * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return st;
}

int hana_target_enable_winner(hana_ctx* c, hana_target* t, int enable) {
    if (!c || !t) return HANA_E_INVALID;
    if (enable && !t->winner) {
        size_t n = (size_t)t->w * t->h;
        if (hipMalloc((void**)&t->winner, 4 * n) != hipSuccess) {
            (void)hipGetLastError();
            return fail(c, HANA_E_NOMEM, "winner map allocation failed");
        }
        HIP_TRY(c, launch_fill32(t->winner, NO_WINNER, n, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));   // draws of either slot may use it next
    } else if (!enable && t->winner) {
        int st = check_status(c);
        if (st) return st;
        (void)sync_all(c);
        (void)hipFree(t->winner);
        t->winner = nullptr;
    }
    return HANA_OK;
}

int hana_target_read_winner(hana_ctx* c, hana_target* t, uint32_t* winner) {
    if (!c || !t || !winner || !t->winner) return HANA_E_INVALID;
    size_t n = (size_t)t->w * t->h;
    int st = check_status(c);
    if (st == HANA_E_DEVICE) return st;
    int r = join(c);
    if (r) return r;
    HIP_TRY(c, hipMemcpyAsync(winner, t->winner, 4 * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return st;
}

int hana_draw(hana_ctx* c, const hana_draw_desc* d) {
    if (!c || !d || !d->mesh || !d->target || !d->uniforms || !valid_kind(d->shader))
        return fail(c, HANA_E_INVALID, "hana_draw: bad descriptor");
    if (d->shadow_map == d->target) return fail(c, HANA_E_INVALID, "hana_draw: shadow map is the target");
    PassSpec s{d->mesh, d->shader, *d->uniforms, d->diffuse, d->normal, d->shadow_map, d->target, 0};
    return run_passes(c, {s});
}

int hana_render_frame(hana_ctx* c, const hana_frame_desc* f) {
    if (!c || !f || !f->mesh || !f->target || !f->uniforms || !valid_kind(f->shader) ||
        (f->enable_shadow && (!f->shadow_map || f->shadow_map == f->target)))
        return fail(c, HANA_E_INVALID, "hana_render_frame: bad descriptor");
    hana_uniforms u = *f->uniforms;
    u.enable_shadow = f->enable_shadow;
    std::vector<PassSpec> specs;
    if (f->enable_shadow) specs.push_back({f->mesh, HANA_SHADER_SHADOW, u, f->diffuse, f->normal, nullptr, f->shadow_map, 0});
    specs.push_back({f->mesh, f->shader, u, f->diffuse, f->normal, f->enable_shadow ? f->shadow_map : nullptr, f->target,
                     f->enable_shadow ? 1 : 0});
    int r = run_passes(c, std::move(specs));
    if (r) return r;
    if (f->enable_shadow) {
        // shdaow_map->renderbuffer_clear_color(Black) / clear_depth(FLT_MAX)
        // after the main pass (SR/scene.h:94-98); alpha is written as 255
        f->shadow_map->pending_clear = true;
        f->shadow_map->clear_rgba = 0xFF000000u;
        f->shadow_map->clear_depth = 3.40282347e+38f;
    }
    return HANA_OK;
}

// ---- batches ----

int hana_batch_create(hana_ctx* c, hana_batch** out) {
    if (!c || !out) return HANA_E_INVALID;
    *out = new hana_batch();
    return HANA_OK;
}

int hana_batch_add_frame(hana_ctx* c, hana_batch* b, const hana_frame_desc* f) {
    if (!c || !b || !f || !f->mesh || !f->target || !f->uniforms || !valid_kind(f->shader) ||
        (f->enable_shadow && (!f->shadow_map || f->shadow_map == f->target)))
        return fail(c, HANA_E_INVALID, "hana_batch_add_frame: bad descriptor");
    if (b->finalized) return fail(c, HANA_E_INVALID, "hana_batch_add_frame: batch already finalized");
    hana_uniforms u = *f->uniforms;
    u.enable_shadow = f->enable_shadow;
    b->frame_first.push_back(b->specs.size());
    if (f->enable_shadow) b->specs.push_back({f->mesh, HANA_SHADER_SHADOW, u, f->diffuse, f->normal, nullptr, f->shadow_map, 0});
    b->specs.push_back({f->mesh, f->shader, u, f->diffuse, f->normal, f->enable_shadow ? f->shadow_map : nullptr,
                        f->target, 1});
    return HANA_OK;
}

// Every frame of a batch clears its target and shadow map first (the state
// DrawModel::draw runs in after main.cpp's per-frame clear).
static void batch_clear_flags(hana_batch_part& part) {
    for (size_t i = 0; i < part.plan.descs.size(); i++) {
        PassDesc& d = part.plan.descs[i];
        d.clear = 1;
        d.clear_rgba = 0xFF000000u;
        d.clear_depth = 3.40282347e+38f;
    }
}

// The sequence itself clears every target it draws, so a clear still pending
// from target creation or hana_target_clear is void: left pending, it would
// wipe the rendered image at the next read.
static void batch_drop_pending(hana_batch* b) {
    for (auto& s : b->specs) s.target->pending_clear = false;
}

// Each frame's first work estimate: faces plus an eighth of the pixels drawn,
// over its passes.
static std::vector<double> frame_weights(const hana_batch* b) {
    size_t nf = b->frame_first.size();
    std::vector<double> w(nf, 0.0);
    for (size_t k = 0; k < nf; k++) {
        size_t e = k + 1 < nf ? b->frame_first[k + 1] : b->specs.size();
        for (size_t i = b->frame_first[k]; i < e; i++)
            w[k] += (double)b->specs[i].mesh->F + (double)b->specs[i].target->w * b->specs[i].target->h / 8.0;
    }
    return w;
}

// The end frame of each part when the frames, in order, are split into up to
// b->tuning.parts parts of about equal estimated work w (one value per frame),
// at most one part per frame.  Four parts measured best on 256 mixed scenes
// (2.10 ms per batch, against 2.20 for two and 2.3-2.4 for six or eight): a
// process gets four hardware queues, and more streams than that share them.
static std::vector<size_t> split_ends(const hana_batch* b, const std::vector<double>& w) {
    size_t nf = b->frame_first.size();
    size_t np = std::min(nf, std::min((size_t)b->tuning.parts, kMaxBatchParts));
    double total = 0.0;
    for (double x : w) total += x;
    // part p ends at the first frame where the running work reaches (p+1)/np
    // of the total, leaving at least one frame for each later part
    std::vector<size_t> ends(np);
    size_t k = 0;
    double run = 0.0;
    for (size_t p = 0; p < np; p++) {
        if (p + 1 == np) {
            k = nf;
        } else {
            do {
                run += w[k++];
            } while (k < nf - (np - 1 - p) && run < total * (double)(p + 1) / (double)np);
        }
        ends[p] = k;
    }
    return ends;
}

// Replaces the batch's parts by those ending at `ends` (unsized: no plan yet).
static void batch_split(hana_batch* b, const std::vector<size_t>& ends) {
    size_t nf = b->frame_first.size();
    b->parts.clear();
    b->parts.resize(ends.size());
    size_t k0 = 0;
    for (size_t p = 0; p < ends.size(); p++) {
        size_t k = ends[p];
        size_t s0 = b->frame_first[k0], s1 = k < nf ? b->frame_first[k] : b->specs.size();
        b->parts[p].specs.assign(b->specs.begin() + s0, b->specs.begin() + s1);
        b->parts[p].arena = &b->arenas[p];
        b->parts[p].f0 = k0;
        b->parts[p].f1 = k;
        k0 = k;
    }
}

// Part p of a batch runs on the stream of draw slot p (part 0 on the
// context stream), and parts beyond the slots on the batch's own streams, so
// that a context and its batches use as few streams as possible.
static hipStream_t part_stream(hana_ctx* c, hana_batch* b, size_t p) {
    return p < (size_t)c->nslots ? slot_stream(c, (int)p) : b->extra[p - c->nslots];
}

// Enqueues every part on its stream, forked from and joined back into the
// context stream (with stage timing on, all run on the context stream so
// that stage times do not overlap).
static int batch_enqueue(hana_ctx* c, hana_batch* b, bool timed) {
    int r;
    if (b->parts.size() == 1 || (timed && c->timing)) {
        for (auto& part : b->parts)
            if ((r = enqueue_sequence(c, part.plan, *part.arena, c->stream, timed))) return r;
        return HANA_OK;
    }
    HIP_TRY(c, hipEventRecord(b->fork, c->stream));
    for (size_t p = 1; p < b->parts.size(); p++) {
        HIP_TRY(c, hipStreamWaitEvent(part_stream(c, b, p), b->fork, 0));
    }
    for (size_t p = 0; p < b->parts.size(); p++) {
        hipStream_t s = part_stream(c, b, p);
        if ((r = enqueue_sequence(c, b->parts[p].plan, *b->parts[p].arena, s, timed))) return r;
    }
    for (size_t p = 1; p < b->parts.size(); p++) {
        hipStream_t s = part_stream(c, b, p);
        HIP_TRY(c, hipEventRecord(b->joined[p - 1], s));
        HIP_TRY(c, hipStreamWaitEvent(c->stream, b->joined[p - 1], 0));
    }
    return HANA_OK;
}

// Sizes a part's workspace exactly with trial runs.
static int batch_size_part(hana_ctx* c, hana_batch* b, hana_batch_part& part) {
    int r;
    if ((r = build_plan(c, part.specs, part.plan, b->tuning))) return r;
    uint32_t min_bins = 0, min_polys = 0, min_tris = 0;
    Arena& A = *part.arena;
    for (int attempt = 0; attempt < 3; attempt++) {
        if ((r = bind_plan(c, part.plan, A, min_bins, min_polys, min_tris, false, c->stream))) return r;
        batch_clear_flags(part);
        HIP_TRY(c, hipMemcpy(A.tables.p, part.plan.descs.data(), sizeof(PassDesc) * part.plan.descs.size(),
                             hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(A.tables.as<char>() + part.plan.uni_off, part.plan.unis.data(),
                             sizeof(DevUniforms) * part.plan.unis.size(), hipMemcpyHostToDevice));
        if (!part.plan.tile_pass.empty())
            HIP_TRY(c, hipMemcpy(A.tables.as<char>() + part.plan.tp_off, part.plan.tile_pass.data(),
                                 sizeof(uint2) * part.plan.tile_pass.size(), hipMemcpyHostToDevice));
        // trial run: sizes the bins and the clip pool exactly
        if ((r = enqueue_sequence(c, part.plan, A, c->stream))) return r;
        uint32_t st[STATUS_WORDS];
        HIP_TRY(c, hipMemcpyAsync(st, A.status.p, sizeof st, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        batch_drop_pending(b);
        if (!st[STATUS_FLAGS]) return HANA_OK;
        HIP_TRY(c, launch_fill32(A.status.as<uint32_t>() + STATUS_SEQ_WORDS, 0u, STATUS_WORDS - STATUS_SEQ_WORDS,
                                 c->stream));
        if (st[STATUS_FLAGS] & FLAG_QUEUE_STALL)
            return fail(c, HANA_E_DEVICE, "hana_batch_finalize: the persistent raster's wait timed out");
        if (!(st[STATUS_FLAGS] & FLAG_CAPACITY)) {
            if (st[STATUS_FLAGS] & FLAG_PAIR_LIMIT)
                return fail(c, HANA_E_NOMEM, "hana_batch_finalize: a part covers more than 4,294,967,280 "
                                             "(triangle, tile) pairs; split the batch");
            return HANA_OK;   // a polygon over POLY_MAX vertices is dropped at every run, like a per-call draw
        }
        min_bins = st[STATUS_BINS_NEEDED] + st[STATUS_BINS_NEEDED] / 8 + 1024;
        min_polys = st[STATUS_POLYS] + st[STATUS_POLYS] / 8 + 1024;
        min_tris = st[STATUS_TRIS] + st[STATUS_TRIS] / 8 + 1024;
    }
    return fail(c, HANA_E_OVERFLOW, "hana_batch_finalize: capacity not reached");
}

static void batch_drop_graph(hana_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Splits, sizes and (with use_graph) captures the batch for b->tuning.
static int batch_prepare(hana_ctx* c, hana_batch* b, bool use_graph) {
    int r;
    batch_drop_graph(b);
    std::vector<double> w = frame_weights(b);
    batch_split(b, split_ends(b, w));
    for (auto& part : b->parts)
        if ((r = batch_size_part(c, b, part))) return r;
    if (b->parts.size() > 1) {
        // One rebalancing step.  The parts run concurrently and the batch ends
        // with its slowest one; with the first estimates the parts of 256
        // mixed scenes ended up to 0.27 ms apart.  Each part is timed alone
        // (a run to warm up, then a timed one: the sequence clears what it
        // draws, so the runs leave nothing behind), its frames' estimates are
        // scaled by its measured time per unit of estimate, and the frames
        // are split again (unless a part measured no time); if any part's frames
        // changed, every part is sized again.
        std::vector<double> cost(w.size(), 0.0);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIP_TRY(c, hipEventCreate(&e0));
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            return fail(c, HANA_E_DEVICE, "hipEventCreate");
        }
        r = HANA_OK;
        bool measured = true;   // every part took a positive time
        for (auto& part : b->parts) {
            float ms = 0.0f;
            for (int rep = 0; rep < 2 && !r; rep++) {
                if (hipEventRecord(e0, c->stream) != hipSuccess) r = fail(c, HANA_E_DEVICE, "hipEventRecord");
                if (!r) r = enqueue_sequence(c, part.plan, *part.arena, c->stream, false);
                if (!r && hipEventRecord(e1, c->stream) != hipSuccess) r = fail(c, HANA_E_DEVICE, "hipEventRecord");
                if (!r && hipEventSynchronize(e1) != hipSuccess) r = fail(c, HANA_E_DEVICE, "hipEventSynchronize");
                if (!r && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) r = fail(c, HANA_E_DEVICE, "hipEventElapsedTime");
            }
            double W = 0.0;
            for (size_t k = part.f0; k < part.f1; k++) W += w[k];
            measured = measured && ms > 0.0f && W > 0.0;
            for (size_t k = part.f0; k < part.f1; k++) cost[k] = W > 0.0 ? w[k] * (double)ms / W : w[k];
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        batch_drop_pending(b);
        if (r) return r;
        std::vector<size_t> ends = split_ends(b, cost), before;
        for (auto& part : b->parts) before.push_back(part.f1);
        if (measured && ends != before) {
            batch_split(b, ends);
            for (auto& part : b->parts)
                if ((r = batch_size_part(c, b, part))) return r;
        }
    }
    if (b->parts.size() > 1 && !b->fork) HIP_TRY(c, hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    while (b->joined.size() + 1 < b->parts.size()) {
        hipEvent_t e = nullptr;
        HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->joined.push_back(e);
    }
    while (b->extra.size() + c->nslots < b->parts.size()) {
        hipStream_t s = nullptr;
        HIP_TRY(c, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        b->extra.push_back(s);
    }
    b->use_graph = use_graph;
    if (b->use_graph) {
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int er = batch_enqueue(c, b, false);
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(c->stream, &g);
        if (er) {
            if (g) (void)hipGraphDestroy(g);
            return er;
        }
        if (ec != hipSuccess) return hip_fail(c, ec, "hipStreamEndCapture");
        b->graph = g;
        HIP_TRY(c, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    return HANA_OK;
}

static int batch_check_new(hana_ctx* c, hana_batch* b, const char* what) {
    if (b->specs.empty()) return fail(c, HANA_E_INVALID, std::string(what) + ": empty batch");
    if (b->finalized) return fail(c, HANA_E_INVALID, std::string(what) + ": batch already finalized");
    for (size_t i = 0; i < b->specs.size(); i++)
        for (size_t j = i + 1; j < b->specs.size(); j++)
            if (b->specs[i].target == b->specs[j].target)
                return fail(c, HANA_E_INVALID, std::string(what) + ": frames of a batch need distinct targets and shadow maps");
    return HANA_OK;
}

int hana_batch_finalize(hana_ctx* c, hana_batch* b, int use_graph) {
    if (!c || !b) return HANA_E_INVALID;
    int r;
    if ((r = batch_check_new(c, b, "hana_batch_finalize"))) return r;
    if ((r = check_status(c)) || (r = join(c))) return r;
    if ((r = batch_prepare(c, b, use_graph != 0))) return r;
    b->finalized = true;
    return HANA_OK;
}

// One run of a prepared batch on the context stream.
static int batch_launch(hana_ctx* c, hana_batch* b) {
    batch_drop_pending(b);
    int r = join(c);
    if (r) return r;
    c->next_slot = 0;   // the next draw follows the sequence on the context stream
    if (b->use_graph) HIP_TRY(c, hipGraphLaunch(b->exec, c->stream));
    else if ((r = batch_enqueue(c, b, true))) return r;
    // every part has joined the context stream: draws on other slots that
    // touch the batch's targets wait for it
    HIP_TRY(c, hipEventRecord(c->slot_done[0], c->stream));
    std::vector<hana_target*> t;
    for (const PassSpec& s : b->specs)
        for (hana_target* x : {s.target, s.shadow})
            if (x) t.push_back(x);
    note_pending(c, 0, t);
    return HANA_OK;
}

// Mean time of `reps` runs after one untimed run, with events on the context
// stream (every part joins it).
static int batch_time(hana_ctx* c, hana_batch* b, int reps, double* ms) {
    int r;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(c, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(c, HANA_E_DEVICE, "hipEventCreate");
    }
    float t = 0.f;
    r = batch_launch(c, b);
    if (!r) r = sync_all(c);
    if (!r && hipEventRecord(e0, c->stream) != hipSuccess) r = fail(c, HANA_E_DEVICE, "hipEventRecord");
    for (int i = 0; i < reps && !r; i++) r = batch_launch(c, b);
    if (!r && hipEventRecord(e1, c->stream) != hipSuccess) r = fail(c, HANA_E_DEVICE, "hipEventRecord");
    if (!r) r = sync_all(c);
    if (!r && hipEventElapsedTime(&t, e0, e1) != hipSuccess) r = fail(c, HANA_E_DEVICE, "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / std::max(reps, 1);
    return r;
}


static void tuning_out(const Tuning& t, double ms, hana_tuning* out) {
    std::memset(out, 0, sizeof *out);
    out->tile_size = tile_edge(t.tile_shift);
    out->raster_wgs_per_cu = t.raster_wgs_per_cu;
    out->parts = t.parts;
    out->raster_waves = t.raster_waves;
    out->ms = ms;
}

int hana_gpu_set_tuning(hana_ctx* c, const hana_tuning* t) {
    if (!c) return HANA_E_INVALID;
    Tuning tu;
    int r = tuning_from(c, t, tu);
    if (r) return r;
    c->tuning = tu;
    return HANA_OK;
}

int hana_batch_set_tuning(hana_ctx* c, hana_batch* b, const hana_tuning* t) {
    if (!c || !b) return HANA_E_INVALID;
    if (b->finalized) return fail(c, HANA_E_INVALID, "hana_batch_set_tuning: batch already finalized");
    return tuning_from(c, t, b->tuning);
}

int hana_batch_get_tuning(hana_ctx* c, const hana_batch* b, hana_tuning* out) {
    if (!c || !b || !out) return HANA_E_INVALID;
    tuning_out(b->tuning, b->tuned_ms, out);
    return HANA_OK;
}

// The candidates, one dimension at a time from the batch's current tuning:
// the tile size, then the raster form (per-tile workgroups or a persistent
// grid of 2, 4 or 8 workgroups per CU), then the waves per raster workgroup,
// then the number of parts; each keeps the best value found so far for the
// others.  A candidate replaces the best only if it is faster by more than
// kAutotuneMargin, so that run-to-run noise between near-equal tunings does
// not change the choice.
static constexpr double kAutotuneMargin = 0.02;

int hana_batch_autotune(hana_ctx* c, hana_batch* b, int use_graph, int reps, hana_tuning* chosen) {
    if (!c || !b || reps < 1) return HANA_E_INVALID;
    int r;
    if ((r = batch_check_new(c, b, "hana_batch_autotune"))) return r;
    if ((r = check_status(c)) || (r = join(c))) return r;
    Tuning best = b->tuning;
    double best_ms = 0.0;
    auto measure = [&](const Tuning& t, double* ms) {
        b->tuning = t;
        int e = batch_prepare(c, b, use_graph != 0);
        return e ? e : batch_time(c, b, reps, ms);
    };
    if ((r = measure(best, &best_ms))) return r;
    const int shifts[] = {4, 5}, wgs[] = {0, 2, 4, 8}, waves[] = {4, 2}, parts[] = {1, 2, 4};
    const int* dims[4] = {shifts, wgs, waves, parts};
    const int counts[4] = {2, 4, 2, 3};
    for (int dim = 0; dim < 4; dim++) {
        const int* vals = dims[dim];
        int nv = counts[dim];
        Tuning base = best;
        for (int i = 0; i < nv; i++) {
            Tuning t = base;
            int* fields[4] = {&t.tile_shift, &t.raster_wgs_per_cu, &t.raster_waves, &t.parts};
            const int curs[4] = {base.tile_shift, base.raster_wgs_per_cu, base.raster_waves, base.parts};
            int& field = *fields[dim];
            const int cur = curs[dim];
            if (vals[i] == cur) continue;
            field = vals[i];
            double ms = 0.0;
            if ((r = measure(t, &ms))) return r;
            if (ms < best_ms * (1.0 - kAutotuneMargin)) {
                best = t;
                best_ms = ms;
            }
        }
    }
    b->tuning = best;
    if ((r = batch_prepare(c, b, use_graph != 0))) return r;
    b->tuned_ms = best_ms;
    b->finalized = true;
    if (chosen) tuning_out(best, best_ms, chosen);
    return HANA_OK;
}

int hana_batch_run(hana_ctx* c, hana_batch* b) {
    if (!c || !b || !b->finalized) return fail(c, HANA_E_INVALID, "hana_batch_run: batch not finalized");
    // per-call draws before the batch are complete (replayed if need be)
    // before the batch draws over their targets
    int r = check_status(c);
    if (r) return r;
    return batch_launch(c, b);
}

int hana_batch_destroy(hana_ctx* c, hana_batch* b) {
    if (!c || !b) return HANA_E_INVALID;
    (void)sync_all(c);
    batch_drop_graph(b);
    for (auto& a : b->arenas) free_arena(a);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (hipEvent_t e : b->joined) (void)hipEventDestroy(e);
    for (hipStream_t s : b->extra) (void)hipStreamDestroy(s);
    delete b;
    return HANA_OK;
}

}  // extern "C"
