// hana_noalloc.cpp -- hana_render_frame_noalloc: one DrawModel::draw frame
// (SR/scene.h:53-99) whose launch-sequence workspace lives in a fixed device
// array instead of buffers the library allocates, for callers that cannot
// allocate (code under stream capture, serving loops that must not call
// hipMalloc).
//
// The array, `scratch`, is kNoAllocScratchBytes of file-scope device memory:
// every GPU the library runs on holds its own copy until the process ends.
// Its device address is taken once per device, at context creation, with
// hipGetSymbolAddress, and only that address is used from the host.  Every
// call on a GPU shares that GPU's array: calls that overlap on one GPU (two
// contexts, or two host threads) race.  The sticky status words are cleared
// by the first call after each status check; everything else a sequence
// needs zeroed, its first kernel zeroes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/hana_gpu.h"
#include "hana_hostmath.h"
#include "hana_internal.h"
#include "hana_launch.h"
#include "hana_types.h"

using namespace hana;
using namespace hana::rt;

namespace {

static __device__ __align__(256) unsigned char scratch[hana::rt::kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Bump allocation of 256-byte aligned pieces of the scratch array.
struct Carver {
    char* base;
    size_t off = 0, cap;
    template <class T>
    T* take(size_t count) {
        size_t bytes = (sizeof(T) * std::max<size_t>(count, 1) + 255) & ~(size_t)255;
        if (off + bytes > cap) {
            off = cap + 1;   // marks the layout as too large
            return nullptr;
        }
        T* p = reinterpret_cast<T*>(base + off);
        off += bytes;
        return p;
    }
    bool fits() const { return off <= cap; }
};

}  // namespace

namespace hana {
namespace rt {

int noalloc_init(hana_ctx* c) {
    if (c->device < 0 || c->device >= kMaxDevices) return HANA_E_DEVICE;
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[c->device]) {
        // the device is current: create_context selected it
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return HANA_E_DEVICE;
        }
        g_scratch[c->device] = p;
    }
    c->scratch = g_scratch[c->device];
    return HANA_OK;
}

}  // namespace rt
}  // namespace hana

extern "C" int hana_render_frame_noalloc(hana_ctx* c, const hana_frame_desc* f) {
    if (!c || !f || !f->mesh || !f->target || !f->uniforms || !valid_kind(f->shader) ||
        (f->enable_shadow && (!f->shadow_map || f->shadow_map == f->target)))
        return fail(c, HANA_E_INVALID, "hana_render_frame_noalloc: bad descriptor");
    if (!c->scratch) return fail(c, HANA_E_DEVICE, "hana_render_frame_noalloc: no scratch address on this device");
    if (c->status_pending)
        return fail(c, HANA_E_INVALID,
                    "hana_render_frame_noalloc: per-call draws are pending; wait for them (hana_gpu_sync) first");
    if (c->timing || c->stats_enabled)
        return fail(c, HANA_E_INVALID, "hana_render_frame_noalloc: not with stage timing or statistics on");

    // the frame's passes, as hana_render_frame orders them
    hana_uniforms u = *f->uniforms;
    u.enable_shadow = f->enable_shadow;
    PassSpec specs[2];
    int n = 0;
    if (f->enable_shadow) specs[n++] = {f->mesh, HANA_SHADER_SHADOW, u, f->diffuse, f->normal, nullptr, f->shadow_map, 0};
    specs[n++] = {f->mesh, f->shader, u, f->diffuse, f->normal, f->enable_shadow ? f->shadow_map : nullptr, f->target,
                  f->enable_shadow ? 1 : 0};

    PassDesc descs[2];
    DevUniforms unis[2];
    const int tshift = c->tuning.tile_shift, tedge = tile_edge(tshift);
    const size_t vchunk = (size_t)BLOCK * VPT, fchunk = (size_t)BLOCK * FPT;
    size_t voff[2], vsoff[2], foff[2], tv = 0, tvs = 0, tf = 0;
    int vb = 0, fb = 0, tile0 = 0, tb[2] = {0, 0}, group_first[2] = {0, 0}, group_count[2] = {0, 0};
    bool small_tris = true;
    for (int i = 0; i < n; i++) {
        const PassSpec& s = specs[i];
        PassDesc& d = descs[i];
        std::memset(&d, 0, sizeof d);
        derive_dev_uniforms(s.u, unis[i]);
        hana_mesh* m = s.mesh;
        size_t V = (size_t)m->V;
        d.nverts = m->V;
        d.nfaces = m->F;
        d.kind = s.kind;
        d.nvary = (s.kind == SHADER_GROUND || s.kind == SHADER_TOON || s.kind == SHADER_SHADOW) ? 1
                  : (s.kind == SHADER_TEXTURE ? 2 : 8);
        d.px = m->soa; d.py = m->soa + V; d.pz = m->soa + 2 * V;
        d.nx = m->soa + 3 * V; d.ny = m->soa + 4 * V; d.nz = m->soa + 5 * V;
        d.tu = m->soa + 6 * V; d.tv = m->soa + 7 * V;
        d.idx = m->idx;
        if (s.diffuse) { d.diffuse = s.diffuse->texels; d.dw = s.diffuse->w; d.dh = s.diffuse->h; }
        if (s.normal) { d.normal = s.normal->texels; d.nw = s.normal->w; d.nh = s.normal->h; }
        if (s.shadow) { d.shadow = s.shadow->color; d.shadow_depth = s.shadow->depth; d.sw = s.shadow->w; d.sh = s.shadow->h; }
        hana_target* t = s.target;
        d.color = t->color; d.depth = t->depth; d.winner = t->winner;
        d.width = t->w; d.height = t->h;
        d.tiles_x = (t->w + tedge - 1) / tedge;
        d.tiles_y = (t->h + tedge - 1) / tedge;
        d.share_small = (int64_t)t->w * t->h > 4 * (int64_t)m->F;
        if (d.share_small) small_tris = false;
        size_t vpad = (V + vchunk - 1) / vchunk * vchunk;
        voff[i] = tv; vsoff[i] = tvs; foff[i] = tf;
        tv += vpad;
        tvs += vpad * (size_t)d.nvary;
        tf += (size_t)m->F;
        d.vblock0 = vb;
        vb += (int)(vpad / vchunk);
        d.fblock0 = fb;
        fb += (int)(((size_t)m->F + fchunk - 1) / fchunk);
        d.tile0 = tile0;
        tile0 += d.tiles_x * d.tiles_y;
        d.tblock0 = tb[s.group];
        tb[s.group] += d.tiles_x * d.tiles_y;
        if (!group_count[s.group]) group_first[s.group] = i;
        group_count[s.group]++;
        // clears fuse into the first pass writing the target
        if (t->pending_clear) {
            d.clear = 1;
            d.clear_rgba = t->clear_rgba;
            d.clear_depth = t->clear_depth;
        }
    }

    // the workspace, carved from the scratch array
    Carver cv{static_cast<char*>(c->scratch), 0, kNoAllocScratchBytes};
    const size_t ntiles = (size_t)tile0 + 1, nsegs = (size_t)std::max(fb, 1);
    const uint32_t seg_stride = small_tris ? 512u : 1024u;
    const uint64_t poly_cap = std::min<uint64_t>(tf / 256 + 1024, POOL_CAP_MAX);
    const uint64_t seg_region = (uint64_t)fb * BLOCK * FPT;
    const uint64_t tris_cap = seg_region + 8 * poly_cap;
    PassDesc* d_desc = cv.take<PassDesc>(2);
    DevUniforms* d_uni = cv.take<DevUniforms>(2);
    Workspace ws{};
    ws.status = cv.take<uint32_t>(STATUS_WORDS);
    ws.qstate = cv.take<uint32_t>(2);
    ws.tile_count = cv.take<uint32_t>(ntiles);
    ws.tile_offset = cv.take<uint32_t>(ntiles);
    ws.tile_cursor = cv.take<uint32_t>(ntiles);
    ws.scan_part = cv.take<uint64_t>((size_t)scan_chunks(tile0));
    float4* vr = cv.take<float4>(tv);
    float* vs = cv.take<float>(tvs);
    uint8_t* code = cv.take<uint8_t>(tv);
    int32_t* pof = cv.take<int32_t>(tf);
    ws.seg_count = cv.take<uint32_t>(nsegs);
    ws.seg_tiles = cv.take<uint2>(nsegs * seg_stride);
    ws.seg_wide = cv.take<uint16_t>(nsegs * BLOCK * FPT);
    ws.seg_info = cv.take<uint32_t>(nsegs);
    ws.clip_queue = cv.take<uint64_t>(std::max<size_t>(tf, 1024));
    ws.pool_vr = cv.take<float4>(POLY_MAX * poly_cap);
    ws.pool_vs = cv.take<float>(POLY_MAX * POLY_VARY * poly_cap);
    ws.pool_face = cv.take<int32_t>(poly_cap);
    ws.pool_pass = cv.take<int32_t>(poly_cap);
    ws.pool_n = cv.take<int32_t>(poly_cap);
    ws.tris = cv.take<uint4>(tris_cap);
    // the bins take what is left, at least four entries per face
    size_t left = cv.fits() ? (kNoAllocScratchBytes - cv.off) / sizeof(uint32_t) : 0;
    size_t bins_cap = std::min<size_t>(left > 64 ? left - 64 : 0, MAX_PAIRS);
    ws.bins = bins_cap ? cv.take<uint32_t>(bins_cap) : nullptr;
    if (!cv.fits() || !ws.bins || bins_cap < 4 * tf + 2 * ntiles || tris_cap > 0xFFFFFFF0ull)
        return fail(c, HANA_E_NOMEM,
                    "hana_render_frame_noalloc: the frame needs more workspace than the " +
                        std::to_string(kNoAllocScratchBytes >> 20) + " MiB no-alloc scratch holds; use hana_render_frame");
    ws.ntiles = tile0;
    ws.bin_cap = (uint32_t)bins_cap;
    ws.poly_cap = (uint32_t)poly_cap;
    ws.clipq_cap = (uint32_t)std::max<size_t>(tf, 1024);
    ws.tris_cap = (uint32_t)tris_cap;
    ws.seg_region = (uint32_t)seg_region;
    ws.seg_stride = seg_stride;
    ws.status_out = c->noalloc_status_dev;
    ws.q_items = (uint32_t)tile0;
    ws.q_g0 = (uint32_t)tb[0];
    ws.q_dep_from = n == 2 ? (uint32_t)descs[1].tile0 : (uint32_t)tile0;
    ws.tile_shift = tshift;
    ws.raster_waves = c->tuning.raster_waves;
    for (int i = 0; i < n; i++) {
        descs[i].vr = vr + voff[i];
        descs[i].vs = vs + vsoff[i];
        descs[i].code = code + voff[i];
        descs[i].poly_of_face = pof + foff[i];
        descs[i].u = d_uni + i;
        descs[i].stats = nullptr;
    }

    // enqueue on the context stream after every draw in flight on the others
    int r = join(c);
    if (r) return r;
    hipStream_t s = c->stream;
#define NA_TRY(expr)                                                         \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) return hip_fail(c, e_, "hana_render_frame_noalloc: " #expr); \
    } while (0)
    // the first call after a status check starts the sticky words from zero
    if (!c->noalloc_pending) NA_TRY(launch_fill32(ws.status, 0u, STATUS_WORDS, s));
    NA_TRY(hipMemcpyAsync(d_desc, descs, sizeof(PassDesc) * n, hipMemcpyHostToDevice, s));
    NA_TRY(hipMemcpyAsync(d_uni, unis, sizeof(DevUniforms) * n, hipMemcpyHostToDevice, s));
    NA_TRY(launch_vertex(d_desc, n, ws, vb, s));
    NA_TRY(launch_setup(d_desc, n, ws, fb, small_tris, s));
    NA_TRY(launch_clip(d_desc, ws, s));
    NA_TRY(launch_scan(ws, s));
    NA_TRY(launch_bin(d_desc, ws, fb, small_tris, s));
    if (c->tuning.raster_wgs_per_cu > 0) {
        NA_TRY(launch_raster_queue(d_desc, n, ws, c->tuning.raster_wgs_per_cu * c->num_cus, s));
    } else {
        for (int g = 0; g < 2; g++)
            if (group_count[g])
                NA_TRY(launch_raster(d_desc + group_first[g], group_count[g], ws, tb[g], g, s));
    }
    NA_TRY(hipEventRecord(c->slot_done[0], s));
#undef NA_TRY
    for (int i = 0; i < n; i++) specs[i].target->pending_clear = false;
    if (f->enable_shadow) {
        // the post-frame shadow-map clear (SR/scene.h:94-98), fused into its next draw
        f->shadow_map->pending_clear = true;
        f->shadow_map->clear_rgba = 0xFF000000u;
        f->shadow_map->clear_depth = 3.40282347e+38f;
    }
    // draws on the other slots wait for this one before they touch anything
    c->noalloc_wait_mask = ((1u << c->nslots) - 1u) & ~1u;
    c->next_slot = 0;
    c->noalloc_pending = true;
    return HANA_OK;
}
