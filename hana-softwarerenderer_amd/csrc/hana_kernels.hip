// hana_kernels.hip -- the gfx950 rasterisation pipeline.
//
// One launch sequence renders any number of passes (a pass is one
// graphics_draw_triangle call, SR/graphics.cpp:378-407):
//
//   k_vertex   VPT vertices per lane: vertex program, clip-space outcode
//              (is_vertex_visible and the 7 plane tests), perspective divide
//              and 1/w                                (SR/IShader.cpp, SR/graphics.cpp:29-49, 132-134, 316-337)
//   k_setup    FPT faces per lane: back-face cull, viewport, bbox, degenerate
//              test of trivially accepted faces; counts their (triangle,
//              tile) pairs and lists them for binning; drops faces that the
//              first plane they cross removes whole; queues the rest for
//              clipping
//   k_clip     clips the queued faces (SR/graphics.cpp:136-161) into polygons,
//              counts and lists their fan triangles
//   k_scan_*   exclusive scan of the per-tile counts (reduce, then apply)
//   k_bin      scatters every listed triangle into its tiles' bins
//   k_raster   one workgroup per 16x16 tile: exact per-pixel coverage and
//              depth (SR/graphics.cpp:222-233, 349-359) resolved in LDS with
//              64-bit (depth, submission order) keys, then one fragment
//              program per covered pixel (SR/graphics.cpp:362-373)
//
// Depth test semantics.  The reference tests `z > stored` per fragment in
// submission order (LEQUAL, later primitive wins ties) and no shader
// discards, so the final pixel is the last fragment in submission order among
// those at the minimum depth.  Each fragment here forms the key
// (order-preserving bits of z, 0xFFFFFFFE - order) and the tile keeps the
// per-pixel minimum with ds_min_u64: the minimum is that same fragment
// whatever order the lanes run in.  NaN depths (which pass any test and make
// the next fragment pass) are resolved exactly by a second sweep.
#include <float.h>
#include <hip/hip_runtime.h>

#include "hana_device.h"
#include "hana_launch.h"

namespace hana {

// ---------------------------------------------------------------------------
// shared helpers
// ---------------------------------------------------------------------------

template <int WHICH>
__device__ __forceinline__ int block0_of(const PassDesc& p) {
    return WHICH == 0 ? p.vblock0 : (WHICH == 1 ? p.fblock0 : p.tblock0);
}
// largest pass index whose first block is <= block (passes sorted by it)
template <int WHICH>
__device__ __forceinline__ int find_pass(const PassDesc* P, int n, int block) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (block0_of<WHICH>(P[mid]) <= block) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}


// Triangle set-up in screen space (SR/graphics.cpp:316-347, SR/maths.cpp:20-25).
struct TriS {
    float ax, ay;        // screen A
    float s0x, s0y;      // C.x - A.x, B.x - A.x
    float s1x, s1y;      // C.y - A.y, B.y - A.y
    float uz;            // cross product z (constant over the triangle)
    float z0, z1, z2;    // screen depths
    int x0, x1, y0, y1;  // pixel loop bounds (inclusive)
};

// Returns false for triangles that can produce no fragment: back-facing
// (is_back_facing, SR/graphics.cpp:172-180), degenerate (|u.z| <= 1e-2, every
// barycentric then (-1,1,1), SR/graphics.cpp:230-232) or with an empty scan.
__device__ __forceinline__ bool tri_setup(float4 a, float4 b, float4 c, int W, int H, TriS& t) {
    float signed_area = a.x * b.y - a.y * b.x + b.x * c.y - b.y * c.x + c.x * a.y - c.y * a.x;
    if (signed_area <= 0) return false;
    float fw = (float)W, fh = (float)H;
    float sx0 = (a.x + 1) * 0.5f * fw, sy0 = (a.y + 1) * 0.5f * fh;
    float sx1 = (b.x + 1) * 0.5f * fw, sy1 = (b.y + 1) * 0.5f * fh;
    float sx2 = (c.x + 1) * 0.5f * fw, sy2 = (c.y + 1) * 0.5f * fh;
    t.z0 = (a.z + 1) * 0.5f;
    t.z1 = (b.z + 1) * 0.5f;
    t.z2 = (c.z + 1) * 0.5f;
    t.ax = sx0;
    t.ay = sy0;
    t.s0x = sx2 - sx0;
    t.s0y = sx1 - sx0;
    t.s1x = sy2 - sy0;
    t.s1y = sy1 - sy0;
    t.uz = t.s0x * t.s1y - t.s0y * t.s1x;
    // (double)|u.z| > 1e-2 is the same test as |u.z| > 0.01f on floats
    if (!(fabsf(t.uz) > 0.01f)) return false;
    float mnx = FLT_MAX, mny = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX;
    float clx = (float)(W - 1), cly = (float)(H - 1);
    mnx = smax(0.f, smin(mnx, sx0)); mxx = smin(clx, smax(mxx, sx0));
    mny = smax(0.f, smin(mny, sy0)); mxy = smin(cly, smax(mxy, sy0));
    mnx = smax(0.f, smin(mnx, sx1)); mxx = smin(clx, smax(mxx, sx1));
    mny = smax(0.f, smin(mny, sy1)); mxy = smin(cly, smax(mxy, sy1));
    mnx = smax(0.f, smin(mnx, sx2)); mxx = smin(clx, smax(mxx, sx2));
    mny = smax(0.f, smin(mny, sy2)); mxy = smin(cly, smax(mxy, sy2));
    // for (P.x = bboxmin.x; P.x <= bboxmax.x; P.x++): start truncates, the
    // bound compares the int as a float
    t.x0 = f2i(mnx);
    t.y0 = f2i(mny);
    if (t.x0 < 0 || t.y0 < 0) return false;
    t.x1 = (mxx >= 0.f) ? (int)floorf(mxx) : -1;
    t.y1 = (mxy >= 0.f) ? (int)floorf(mxy) : -1;
    return t.x1 >= t.x0 && t.y1 >= t.y0;
}

// n / u.z as the hardware's IEEE division sequence computes it when its
// scaling steps are no-ops, from the reciprocal r = cover_recip(u.z) that
// sequence refines (the same operations in the same order, so the same bits):
// q = n r; e = fma(-u.z, q, n); q = fma(e, r, q); e = fma(-u.z, q, n);
// q = fma(e, r, q).  See tri_cover_fast for when that holds.
__device__ __forceinline__ float cover_recip(float uz) {
    float r = __builtin_amdgcn_rcpf(uz);
    float e = fmaf(-uz, r, 1.0f);
    return fmaf(e, r, r);
}
__device__ __forceinline__ float cover_quot(float n, float uz, float r) {
    float q = n * r;
    float e = fmaf(-uz, q, n);
    q = fmaf(e, r, q);
    e = fmaf(-uz, q, n);
    return fmaf(e, r, q);
}

// barycentric + coverage + interpolate_depth at integer pixel (px, py)
// (SR/graphics.cpp:186-194, 222-233, 352-356), the three divisions by u.z
// from one refined reciprocal.  The division sequence (div_scale, rcp, four
// fma, div_fmas, div_fixup) scales its operands only when the quotient, the
// reciprocal or the numerator is near the end of the exponent range, and its
// fixup changes the result only for zero, infinite or NaN operands or a
// quotient out of range.  With the triangle's screen coordinates and edges
// within 2^22 and pixels within 2^20, |n| <= 2^47 and 2^-7 < |u.z| <= 2^45,
// so for every numerator with |n| >= 2^-80 neither happens and cover_quot
// gives the division's bits.  A zero numerator gives a zero of either sign,
// which changes neither the coverage test nor the depth (summed from +0).
// Other candidates divide.
__device__ __forceinline__ bool tri_cover_fast(const TriS& t, int px, int py, float& wx, float& wy, float& wz, float& z) {
    float Px = (float)px, Py = (float)py;
    float s0z = t.ax - Px;
    float s1z = t.ay - Py;
    float ux = t.s0y * s1z - s0z * t.s1y;
    float uy = s0z * t.s1x - t.s0x * s1z;
    float sum = ux + uy;
    float ext = fmaxf(fmaxf(fmaxf(fabsf(t.ax), fabsf(t.ay)), fmaxf(fabsf(t.s0x), fabsf(t.s0y))),
                      fmaxf(fabsf(t.s1x), fabsf(t.s1y)));
    float r = cover_recip(t.uz);
    // The shortcut needs every nonzero numerator to be at least 2^-80 in
    // magnitude.  When every input of the triangle is zero or at least 2^-8 in
    // magnitude (and at most 2^22), a nonzero numerator is at least 2^-62:
    // ax and ay are multiples of 2^-31, so s0z and s1z are (their rounding
    // keeps that); a product of a nonzero edge component and a nonzero s0z or
    // s1z is at least 2^-39, so a multiple of 2^-62; their differences and
    // the sum stay multiples of 2^-62.  Such a triangle needs no check per
    // candidate; the others (an input under 2^-8 but not zero: rare) divide.
    // v is zero or at least 2^-8 in magnitude when (bits(v) << 1) - 1 >=
    // (bits(2^-8) << 1) - 1 (zero wraps to the maximum).
    const uint32_t kMin8 = ((uint32_t)(127 - 8) << 24) - 1u;
    uint32_t mt = min(min(min((__float_as_uint(t.ax) << 1) - 1u, (__float_as_uint(t.ay) << 1) - 1u),
                          min((__float_as_uint(t.s0x) << 1) - 1u, (__float_as_uint(t.s0y) << 1) - 1u)),
                      min((__float_as_uint(t.s1x) << 1) - 1u, (__float_as_uint(t.s1y) << 1) - 1u));
    bool fast = ext <= 0x1p22f && mt >= kMin8;
    if (fast) {
        wx = 1.f - cover_quot(sum, t.uz, r);
        wy = cover_quot(uy, t.uz, r);
        wz = cover_quot(ux, t.uz, r);
    } else {
        wx = 1.f - sum / t.uz;
        wy = uy / t.uz;
        wz = ux / t.uz;
    }
    if (wx < 0 || wy < 0 || wz < 0) return false;
    z = 0.0f;
    z += t.z2 * wz;
    z += t.z1 * wy;
    z += t.z0 * wx;
    return true;
}

// order-preserving 32-bit image of a depth (-0 folded onto +0)
__device__ __forceinline__ uint32_t depth_bits(float z) {
    uint32_t u = __float_as_uint(z);
    if ((u & 0x7FFFFFFFu) == 0) u = 0;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ uint64_t init_key(float d) {
    uint32_t hi = (d != d) ? 0xFFFFFFFFu : depth_bits(d);   // NaN: the first fragment passes
    return ((uint64_t)hi << 32) | 0xFFFFFFFFull;
}

__device__ __forceinline__ void stat_add(const PassDesc& P, int which, uint64_t v) {
    if (P.stats && v) atomicAdd((unsigned long long*)&P.stats[which], (unsigned long long)v);
}

// Raises a flag for this sequence and in the sticky word the host reads.
__device__ __forceinline__ void raise_flag(const Workspace& ws, uint32_t f) {
    atomicOr(&ws.status[STATUS_SEQ_FLAGS], f);
    atomicOr(&ws.status[STATUS_FLAGS], f);
}

// clip planes in clip_triangle's order (SR/graphics.cpp:136-161) and their
// inside tests (is_inside_plane, SR/graphics.cpp:29-49)
enum { PL_W, PL_PX, PL_NX, PL_PY, PL_NY, PL_PZ, PL_NZ };

__device__ __forceinline__ bool inside_plane(const float* c, int plane) {
    switch (plane) {
    case PL_W: return c[3] >= 1e-5f;
    case PL_PX: return c[0] <= +c[3];
    case PL_NX: return c[0] >= -c[3];
    case PL_PY: return c[1] <= +c[3];
    case PL_NY: return c[1] >= -c[3];
    case PL_PZ: return c[2] <= +c[3];
    default: return c[2] >= -c[3];
    }
}

// ---------------------------------------------------------------------------
// k_vertex
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(BLOCK) void k_vertex(const PassDesc* __restrict__ passes, int npasses, Workspace ws) {
    int gtid = blockIdx.x * BLOCK + threadIdx.x;
    int gsize = gridDim.x * BLOCK;
    for (int i = gtid; i < ws.ntiles; i += gsize) gst(ws.tile_count + i, 0u);
    if (gtid < STATUS_SEQ_WORDS) gst(ws.status + gtid, 0u);
    if (gtid < 2) gst(ws.qstate + gtid, 0u);

    const PassDesc& P = passes[find_pass<0>(passes, npasses, blockIdx.x)];
    Uniforms u = uniforms_of(P.u);
    int base = (blockIdx.x - P.vblock0) * BLOCK * VPT;
    bool need_n = P.kind != SHADER_SHADOW && P.kind != SHADER_TEXTURE;
    bool need_uv = P.kind != SHADER_SHADOW && P.kind != SHADER_GROUND && P.kind != SHADER_TOON;
    // all loads of the thread's VPT vertices first (coalesced SoA streams)
    F3 pos[VPT], nrm[VPT];
    float tu[VPT], tv[VPT];
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        int v = base + k * BLOCK + threadIdx.x;
        pos[k] = f3(0.f, 0.f, 0.f);
        nrm[k] = f3(0.f, 0.f, 0.f);
        tu[k] = tv[k] = 0.f;
        if (v < P.nverts) {
            pos[k] = f3(gld(P.px + v), gld(P.py + v), gld(P.pz + v));
            if (need_n) nrm[k] = f3(gld(P.nx + v), gld(P.ny + v), gld(P.nz + v));
            if (need_uv) {
                tu[k] = gld(P.tu + v);
                tv[k] = gld(P.tv + v);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < VPT; k++) {
        int v = base + k * BLOCK + threadIdx.x;
        if (v < P.nverts) {
            float clip[4], vary[8];
            vertex_program(P.kind, u, pos[k], nrm[k], tu[k], tv[k], clip, vary);
            // is_vertex_visible (SR/graphics.cpp:132-134), then the plane tests
            uint32_t code =
                (fabsf(clip[0]) <= clip[3] && fabsf(clip[1]) <= clip[3] && fabsf(clip[2]) <= clip[3]) ? CODE_VISIBLE : 0u;
#pragma unroll
            for (int pl = 0; pl < 7; pl++)
                if (!inside_plane(clip, pl)) code |= 1u << pl;
            gst(P.code + v, (uint8_t)code);
            // clip_pos / clip_pos[3] and 1 / clip_pos[3] (SR/graphics.cpp:317, 336)
            float4 r;
            r.w = 1 / clip[3];
            r.z = clip[2] / clip[3];
            r.y = clip[1] / clip[3];
            r.x = clip[0] / clip[3];
            gst4(P.vr + v, r);
            if (P.nvary == 8) {
                float4* d = reinterpret_cast<float4*>(P.vs + (size_t)v * 8);
                gst4(d, make_float4(vary[0], vary[1], vary[2], vary[3]));
                gst4(d + 1, make_float4(vary[4], vary[5], vary[6], vary[7]));
            } else if (P.nvary == 2) {
                gst2(reinterpret_cast<float2*>(P.vs) + v, make_float2(vary[0], vary[1]));
            } else {
                gst(P.vs + v, vary[0]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// clipping (SR/graphics.cpp:7-161), for faces failing the trivial accept
// ---------------------------------------------------------------------------

__device__ __forceinline__ float intersect_ratio(const float* p, const float* c, int plane) {
    switch (plane) {
    case PL_W: return (p[3] - 1e-5f) / (p[3] - c[3]);
    case PL_PX: return (p[3] - p[0]) / ((p[3] - p[0]) - (c[3] - c[0]));
    case PL_NX: return (p[3] + p[0]) / ((p[3] + p[0]) - (c[3] + c[0]));
    case PL_PY: return (p[3] - p[1]) / ((p[3] - p[1]) - (c[3] - c[1]));
    case PL_NY: return (p[3] + p[1]) / ((p[3] + p[1]) - (c[3] + c[1]));
    case PL_PZ: return (p[3] - p[2]) / ((p[3] - p[2]) - (c[3] - c[2]));
    default: return (p[3] + p[2]) / ((p[3] + p[2]) - (c[3] + c[2]));
    }
}

// Orders the LDS accesses of the lanes of one wave (the wave executes them in
// order; this keeps the compiler from moving them across).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A queued face is clipped by a group of CLIP_GROUP lanes of one wave, lane i
// of the group owning vertex i of the polygon.  Its two polygon buffers
// (clip position then POLY_VARY varyings per vertex) live in LDS.
constexpr int CLIP_GROUP = 16;
constexpr int CLIP_THREADS = 512;
constexpr int CLIP_FACES = CLIP_THREADS / CLIP_GROUP;   // faces per workgroup and round
constexpr int CLIP_NF = 4 + POLY_VARY;
static_assert(POLY_MAX <= CLIP_GROUP, "one lane per polygon vertex");
static_assert(POLY_MAX - 2 <= (1 << FAN_BITS), "fan index fits its bin-entry field");

// The homogeneous divide of a clip-space vertex, as the face loop applies it
// to the clipped vertices (SR/graphics.cpp:317, 336).
__device__ __forceinline__ float4 ndc_of(const float* v) {
    float w = v[3];
    float4 r;
    r.w = 1 / w;
    r.z = v[2] / w;
    r.y = v[1] / w;
    r.x = v[0] / w;
    return r;
}

// ---------------------------------------------------------------------------
// binning: count the (triangle, tile) pairs and list the triangles; after the
// scan, k_bin scatters the listed triangles into the bins
// ---------------------------------------------------------------------------

// Reserves one slot in counters[t] for every lane with `want`, with one atomic
// per distinct t in the wave; returns the lane's slot.  The lanes are grouped
// by t first, without any memory access; then every group's first lane issues
// the group's atomic, and the returned bases are awaited once for all groups
// (awaiting each group's atomic before forming the next group serialises their
// round trips where a wave's tiles are scattered).  Every active lane of the
// wave must call it together (the loop is wave-uniform).
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* counters, int t, bool want) {
    int lane = __lane_id();
    bool pending = want;
    uint64_t mask = __ballot(pending);
    int leader_of = lane;
    uint32_t rank = 0, group = 0;
    while (mask) {
        int leader = __ffsll((unsigned long long)mask) - 1;
        int lt = __shfl(t, leader);
        bool match = pending && t == lt;
        uint64_t same = __ballot(match);
        if (match) {
            leader_of = leader;
            rank = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
            pending = false;
        }
        if (lane == leader) group = (uint32_t)__popcll(same);
        mask &= ~same;
    }
    uint32_t base = 0;
    if (want && leader_of == lane) base = gadd(counters + t, group);
    base = __shfl(base, leader_of);
    return base + rank;
}

// Adds one to counters[t] for every lane with `want`: one atomic per distinct
// t in the wave, none of them waited for.  Wave-uniform like wave_reserve.
__device__ __forceinline__ void wave_count(uint32_t* counters, int t, bool want) {
    bool pending = want;
    uint64_t mask = __ballot(pending);
    int lane = __lane_id();
    while (mask) {
        int leader = __ffsll((unsigned long long)mask) - 1;
        int lt = __shfl(t, leader);
        bool match = pending && t == lt;
        uint64_t same = __ballot(match);
        if (lane == leader) gadd(counters + lt, (uint32_t)__popcll(same));
        if (match) pending = false;
        mask = __ballot(pending);
    }
}

// The tiles of one workgroup's (triangle, tile) pairs in an LDS hash table:
// open addressing over BIN_HASH slots, a bounded probe; a tile that finds no
// slot goes through the global cursors instead.
// k_setup and k_bin size their tables by the sequence: 512 slots when every
// pass has small triangles (a segment of 1024 faces then touches a few dozen
// tiles, and a smaller table is quicker to clear and scan), more otherwise.
constexpr int BIN_PROBES = 32;
constexpr uint32_t BIN_EMPTY = 0xFFFFFFFFu;

template <int BITS>
__device__ __forceinline__ int bin_slot(uint32_t* keys, uint32_t gt, bool insert) {
    constexpr uint32_t SLOTS = 1u << BITS;
    uint32_t h = (gt * 2654435761u) >> (32 - BITS);
    for (int p = 0; p < BIN_PROBES; p++) {
        uint32_t sl = (h + (uint32_t)p) & (SLOTS - 1);
        uint32_t k = keys[sl];
        if (k == gt) return (int)sl;
        if (k == BIN_EMPTY) {
            if (!insert) return -1;
            uint32_t old = atomicCAS(&keys[sl], BIN_EMPTY, gt);
            if (old == BIN_EMPTY || old == gt) return (int)sl;
        }
    }
    return -1;
}

// Counts one (triangle, tile) pair in the workgroup's LDS table, or in the
// global counter when the table has no slot left for the tile.
template <int BITS>
__device__ __forceinline__ void lds_count(uint32_t* keys, uint32_t* cnt, uint32_t* global, uint32_t gt) {
    int sl = bin_slot<BITS>(keys, gt, true);
    if (sl >= 0) atomicAdd(&cnt[sl], 1u);
    else gadd(global + gt, 1u);
}

// A workgroup barrier that orders LDS accesses only: unlike __syncthreads it
// does not wait for the wave's outstanding global loads first, so loads
// issued before it stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Triangles over more tiles than this are counted and binned by a whole wave
// (or workgroup), their tiles spread over the lanes.
constexpr int WIDE_TILES = 16;

// Counts every (triangle, tile) pair of one lane's triangle and writes its
// record (bin entry, pass, first and last tile) for k_bin to ws.tris[slot];
// wave-uniform.  With BITS > 0 the pairs of triangles over at most WIDE_TILES
// tiles go to the workgroup's LDS table (keys, cnt), which the caller adds to
// the global counts; with BITS == 0 they go to the global counts directly,
// one atomic per distinct tile of the wave.
struct SegTables {       // k_setup's LDS tables of one segment
    uint32_t* keys;
    uint32_t* cnt;
    uint16_t* wide;      // segment-local indices of the triangles over more than WIDE_TILES tiles
    uint32_t* nwide;
};

template <int BITS, int TS>
__device__ __forceinline__ void count_tiles(const PassDesc& P, const Workspace& ws, int pass, bool ok, const TriS& t,
                                            uint32_t entry, uint32_t slot, uint32_t local, const SegTables* L) {
    int tx0 = 0, ty0 = 0, tx1 = 0, ty1 = 0, nx = 0, n = 0;
    if (ok) {
        tx0 = t.x0 >> TS;
        ty0 = t.y0 >> TS;
        tx1 = t.x1 >> TS;
        ty1 = t.y1 >> TS;
        nx = tx1 - tx0 + 1;
        n = nx * (ty1 - ty0 + 1);
        gst4(ws.tris + slot, make_uint4(entry, (uint32_t)pass, (uint32_t)tx0 | ((uint32_t)ty0 << 16),
                                        (uint32_t)tx1 | ((uint32_t)ty1 << 16)));
    }
    // a triangle over more than WIDE_TILES tiles is counted by the whole wave
    int tile0 = P.tile0, tiles_x = P.tiles_x;
    if constexpr (BITS > 0) {
        if (n > WIDE_TILES) L->wide[atomicAdd(L->nwide, 1u)] = (uint16_t)local;
    }
    uint64_t wide = __ballot(n > WIDE_TILES);
    while (wide) {
        int src = __ffsll((unsigned long long)wide) - 1;
        wide &= wide - 1;
        int wn = __shfl(n, src), wnx = __shfl(nx, src), wx0 = __shfl(tx0, src), wy0 = __shfl(ty0, src);
        int wt0 = __shfl(tile0, src), wtx = __shfl(tiles_x, src);
        for (int k = __lane_id(); k < wn; k += 64)
            gadd(ws.tile_count + (wt0 + (wy0 + k / wnx) * wtx + wx0 + k % wnx), 1u);
    }
    int ns = n > WIDE_TILES ? 0 : n;
    if constexpr (BITS == 0) {
        for (int k = 0; __ballot(k < ns); k++) {
            bool want = k < ns;
            int gt = want ? tile0 + (ty0 + k / nx) * tiles_x + tx0 + k % nx : 0;
            wave_count(ws.tile_count, gt, want);
        }
    } else {
        for (int k = 0; k < ns; k++)
            lds_count<BITS>(L->keys, L->cnt, ws.tile_count, (uint32_t)(tile0 + (ty0 + k / nx) * tiles_x + tx0 + k % nx));
    }
    if (P.stats) {
        for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
        if (__lane_id() == 0) stat_add(P, ST_ENTRIES, (uint64_t)n);
    }
}

// A slot in an LDS counter for every lane with `want`, one LDS atomic per
// wave; wave-uniform.
__device__ __forceinline__ uint32_t wave_reserve_lds(uint32_t* counter, bool want) {
    uint64_t m = __ballot(want);
    int lane = __lane_id();
    uint32_t base = 0;
    int leader = m ? __ffsll((unsigned long long)m) - 1 : 0;
    if (m && lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// Faces that fail the trivial accept, from the outcodes of their vertices:
// clip_triangle passes a polygon whose vertices are all inside a plane through
// that plane unchanged, and returns no vertex once a plane has all of them
// outside (SR/graphics.cpp:73-110, 150-152).  So a face dies whole if, at the
// first plane some vertex is outside of, every vertex is outside.
__device__ __forceinline__ bool dies_at_first_crossed_plane(uint32_t c0, uint32_t c1, uint32_t c2) {
    uint32_t some_out = (c0 | c1 | c2) & CODE_PLANES;
    if (!some_out) return false;
    uint32_t first = some_out & (0u - some_out);
    return (c0 & c1 & c2 & first) != 0u;
}

// FPT faces per lane.  Trivially accepted faces are set up from the shared
// post-transform vertices and listed in this workgroup's segment of ws.tris
// (BLOCK * FPT records from blockIdx.x * BLOCK * FPT; the count goes to
// ws.seg_count); faces that clipping would remove are dropped; the others are
// queued for k_clip with one global atomic per workgroup.
// Six waves per SIMD: the gathers of all FPT faces in flight take about 80
// VGPRs; at 64 (eight waves) the compiler serialises them and the kernel is
// slower, at 97 (five) it is slower too.
template <int BITS, int TS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_setup(const PassDesc* __restrict__ passes, int npasses, Workspace ws) {
    constexpr int SLOTS = 1 << BITS;
    __shared__ uint64_t s_queue[BLOCK * FPT];
    __shared__ uint32_t s_keys[SLOTS], s_cnt[SLOTS];   // tile counts of the workgroup's triangles
    __shared__ uint16_t s_wide[BLOCK * FPT];
    __shared__ uint32_t s_ntri, s_nq, s_qbase, s_nwide, s_nlist;
    if (threadIdx.x == 0) { s_ntri = 0; s_nq = 0; s_nwide = 0; s_nlist = 0; }
    const SegTables L{s_keys, s_cnt, s_wide, &s_nwide};
    int pass = find_pass<1>(passes, npasses, blockIdx.x);
    const PassDesc& P = passes[pass];
    int base = (blockIdx.x - P.fblock0) * BLOCK * FPT;
    uint32_t seg = (uint32_t)blockIdx.x * (BLOCK * FPT);
    // Lanes past the pass's last face load that face (a pass that owns a
    // workgroup has one), so that no load waits on a branch: the indices of
    // all FPT faces are in flight together, then the outcodes and
    // post-transform vertices of all of them (gathered together, so the
    // vertices do not wait for the trivial-accept test).  They are issued
    // before the tile table is cleared, so that the clear and its barrier
    // overlap them.
    int ids[FPT][3];
    bool valid[FPT];
#pragma unroll
    for (int k = 0; k < FPT; k++) {
        int f = base + k * BLOCK + threadIdx.x;
        valid[k] = f < P.nfaces;
        int fl = valid[k] ? f : P.nfaces - 1;
        ids[k][0] = gld(P.idx + 3 * fl);
        ids[k][1] = gld(P.idx + 3 * fl + 1);
        ids[k][2] = gld(P.idx + 3 * fl + 2);
    }
    uint32_t codes[FPT][3];
    float4 va[FPT], vb[FPT], vc[FPT];
#pragma unroll
    for (int k = 0; k < FPT; k++) {
        for (int j = 0; j < 3; j++) codes[k][j] = gld(P.code + ids[k][j]);
        va[k] = gld4(P.vr + ids[k][0]);
        vb[k] = gld4(P.vr + ids[k][1]);
        vc[k] = gld4(P.vr + ids[k][2]);
    }
    for (int i = threadIdx.x; i < SLOTS; i += BLOCK) {
        s_keys[i] = BIN_EMPTY;
        s_cnt[i] = 0u;
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < FPT; k++)
        for (int j = 0; j < 3; j++) codes[k][j] = valid[k] ? codes[k][j] : 0u;
    bool trivial[FPT];
#pragma unroll
    for (int k = 0; k < FPT; k++) trivial[k] = valid[k] && (codes[k][0] & codes[k][1] & codes[k][2] & CODE_VISIBLE);
#pragma unroll
    for (int k = 0; k < FPT; k++) {
        int f = base + k * BLOCK + threadIdx.x;
        bool clipped = valid[k] && !trivial[k];
        bool clipq = clipped && !dies_at_first_crossed_plane(codes[k][0], codes[k][1], codes[k][2]);
        if (clipped && !clipq) gst(P.poly_of_face + f, -1);
        uint32_t q = wave_reserve_lds(&s_nq, clipq);
        if (clipq) s_queue[q] = ((uint64_t)(uint32_t)pass << 32) | (uint32_t)f;
        if (P.stats) {
            uint64_t nv = __popcll(__ballot(valid[k])), nc = __popcll(__ballot(clipped));
            if (__lane_id() == 0) {
                stat_add(P, ST_FACES, nv);
                stat_add(P, ST_CLIPPED, nc);
            }
        }
        TriS t;
        bool ok = trivial[k] && tri_setup(va[k], vb[k], vc[k], P.width, P.height, t);
        uint32_t slot = wave_reserve_lds(&s_ntri, ok);
        count_tiles<BITS, TS>(P, ws, pass, ok, t, (uint32_t)f, seg + slot, slot, &L);
    }
    __syncthreads();
    // add the segment's counts to the global ones, and hand k_bin the list of
    // its tiles with their counts and the list of its wide triangles
    uint2* list = ws.seg_tiles + (size_t)blockIdx.x * SLOTS;
    for (int i = threadIdx.x; i < SLOTS; i += BLOCK)
        if (s_keys[i] != BIN_EMPTY) {
            gadd(ws.tile_count + s_keys[i], s_cnt[i]);
            gst2(list + atomicAdd(&s_nlist, 1u), make_uint2(s_keys[i], s_cnt[i]));
        }
    uint32_t nwide = s_nwide;
    for (uint32_t i = threadIdx.x; i < nwide; i += BLOCK) gst(ws.seg_wide + (size_t)blockIdx.x * (BLOCK * FPT) + i, s_wide[i]);
    if (threadIdx.x == 0) {
        gst(ws.seg_count + blockIdx.x, s_ntri);
        s_qbase = s_nq ? atomicAdd(&ws.status[STATUS_CLIPQ], s_nq) : 0u;
    }
    __syncthreads();
    if (threadIdx.x == 0) gst(ws.seg_info + blockIdx.x, s_nlist | (nwide << 16));
    for (uint32_t i = threadIdx.x; i < s_nq; i += BLOCK) {
        uint32_t q = s_qbase + i;
        if (q < ws.clipq_cap) gst(ws.clip_queue + q, s_queue[i]);
        else raise_flag(ws, FLAG_CLIPQ_OVERFLOW);
    }
}

// Clips the queued faces, CLIP_FACES per workgroup and round, grid-stride
// over the queue.  Per group of CLIP_GROUP lanes: the vertex program runs on
// the three corners in three lanes (the same program as k_vertex, so the
// same bits); then each of the seven planes of clip_triangle
// (SR/graphics.cpp:136-161) in order, one Sutherland-Hodgman stage
// (SR/graphics.cpp:73-110) computed vertex-parallel: lane i looks at the edge
// (i-1, i) and emits the intersection when the two ends differ in side and
// vertex i when it is inside, at the place the prefix sum of the lanes'
// emissions gives it -- the order and the arithmetic of the sequential loop.
// A stage with every vertex inside would copy the polygon unchanged and is
// skipped.  Pool slots (one per polygon) and the fan triangles' records
// (after the setup segments of ws.tris) are reserved once per workgroup and
// round.
template <int TS>
__global__ __launch_bounds__(CLIP_THREADS) void k_clip(const PassDesc* __restrict__ passes, Workspace ws) {
    __shared__ float buf[CLIP_FACES][2][POLY_MAX][CLIP_NF];
    __shared__ uint32_t s_npoly, s_ntri, s_poly0, s_tri0;
    uint32_t count = min(ws.status[STATUS_CLIPQ], ws.clipq_cap);
    uint32_t clip_cap = ws.tris_cap - ws.seg_region;
    const int tid = threadIdx.x, lane = tid & 63, g = tid / CLIP_GROUP, li = tid % CLIP_GROUP;
    const uint64_t gmask = 0xFFFFull << (lane & ~(CLIP_GROUP - 1));
    for (uint32_t base = blockIdx.x * CLIP_FACES; base < count; base += gridDim.x * CLIP_FACES) {
        if (tid == 0) { s_npoly = 0; s_ntri = 0; }
        uint32_t q = base + (uint32_t)g;
        bool have = q < count;                       // uniform within the group
        int pass = 0, f = 0;
        if (have) {
            uint64_t e = gld(ws.clip_queue + q);
            pass = (int)(e >> 32);
            f = (int)(uint32_t)e;
        }
        const PassDesc& P = passes[pass];
        if (have && li < 3) {
            int v = gld(P.idx + 3 * f + li);
            bool need_n = P.kind != SHADER_SHADOW && P.kind != SHADER_TEXTURE;
            F3 p = f3(gld(P.px + v), gld(P.py + v), gld(P.pz + v));
            F3 nn = need_n ? f3(gld(P.nx + v), gld(P.ny + v), gld(P.nz + v)) : f3(0.f, 0.f, 0.f);
            float clip[4], vary[8];
            for (int c = 0; c < 8; c++) vary[c] = 0.0f;
            vertex_program(P.kind, uniforms_of(P.u), p, nn, gld(P.tu + v), gld(P.tv + v), clip, vary);
            for (int c = 0; c < 4; c++) buf[g][0][li][c] = clip[c];
            for (int c = 0; c < POLY_VARY; c++) buf[g][0][li][4 + c] = vary[c];
        }
        wave_sync();
        int n = have ? 3 : 0, cur = 0;
        bool too_large = false;
        for (int plane = PL_W; plane <= PL_NZ; plane++) {
            bool act = li < n;
            bool cin = act && inside_plane(buf[g][cur][li], plane);
            uint64_t am = __ballot(act) & gmask, im = __ballot(cin) & gmask;
            if (n < 3 || am == im) continue;         // dead, or every vertex inside
            int prev = (li - 1 + n) % n;
            bool pin = act && inside_plane(buf[g][cur][prev], plane);
            int emit = act ? (pin != cin) + (cin ? 1 : 0) : 0;
            int incl = emit;
            for (int d = 1; d < CLIP_GROUP; d <<= 1) {
                int o = __shfl_up(incl, d, CLIP_GROUP);
                if (li >= d) incl += o;
            }
            int m = __shfl(incl, CLIP_GROUP - 1, CLIP_GROUP);
            if (m > POLY_MAX) {                      // only possible through rounding
                too_large = true;
                n = 0;
                continue;
            }
            int o = incl - emit;
            float(*in)[CLIP_NF] = buf[g][cur];
            float(*out)[CLIP_NF] = buf[g][cur ^ 1];
            if (act && pin != cin) {
                float t = intersect_ratio(in[prev], in[li], plane);
                for (int k = 0; k < CLIP_NF; k++) {
                    float fv = in[prev][k];
                    out[o][k] = fv + (in[li][k] - fv) * t;
                }
                o++;
            }
            if (cin)
                for (int k = 0; k < CLIP_NF; k++) out[o][k] = in[li][k];
            n = m < 3 ? 0 : m;
            cur ^= 1;
            wave_sync();
        }
        if (too_large && li == 0) raise_flag(ws, FLAG_POLY_TOO_LARGE);
        float(*poly)[CLIP_NF] = buf[g][cur];
        bool keep = n >= 3;                          // uniform within the group
        // fan triangle (0, li+1, li+2) of the polygon in lane li
        TriS t;
        bool ok = keep && li < n - 2 && tri_setup(ndc_of(poly[0]), ndc_of(poly[li + 1]), ndc_of(poly[li + 2]), P.width,
                                                P.height, t);
        __syncthreads();
        uint32_t lpoly = wave_reserve_lds(&s_npoly, keep && li == 0);
        uint32_t ltri = wave_reserve_lds(&s_ntri, ok);
        __syncthreads();
        if (tid == 0) {
            s_poly0 = s_npoly ? atomicAdd(&ws.status[STATUS_POLYS], s_npoly) : 0u;
            s_tri0 = s_ntri ? atomicAdd(&ws.status[STATUS_TRIS], s_ntri) : 0u;
        }
        __syncthreads();
        uint32_t pid = __shfl(s_poly0 + lpoly, lane & ~(CLIP_GROUP - 1));
        bool reserved = ok;
        if (keep && pid >= ws.poly_cap) {
            if (li == 0) raise_flag(ws, FLAG_POLY_OVERFLOW);
            keep = false;
            ok = false;
        }
        uint32_t slot = s_tri0 + ltri;
        if (reserved && slot >= clip_cap) {
            raise_flag(ws, FLAG_TRIS_OVERFLOW);
            ok = reserved = false;
        }
        // a reserved record of a dropped polygon is written empty (no tiles) for k_bin
        if (reserved && !ok) gst4(ws.tris + (ws.seg_region + slot), make_uint4(0u, 0u, 1u, 0u));
        if (keep && li < n) {
            gst4(ws.pool_vr + ((size_t)pid * POLY_MAX + li), ndc_of(poly[li]));
            for (int c = 0; c < POLY_VARY; c++) gst(ws.pool_vs + (((size_t)pid * POLY_MAX + li) * POLY_VARY + c), poly[li][4 + c]);
        }
        if (have && li == 0) {
            if (keep) {
                gst(ws.pool_face + pid, f);
                gst(ws.pool_pass + pid, pass);
                gst(ws.pool_n + pid, n);
                stat_add(P, ST_POLYS, 1);
            }
            gst(P.poly_of_face + f, keep ? (int)pid : -1);
        }
        count_tiles<0, TS>(P, ws, pass, ok, t, CLIPPED_BIT | (pid << FAN_BITS) | (uint32_t)li, ws.seg_region + slot,
                           0u, nullptr);
        __syncthreads();
    }
}

// Scatters one listed triangle into the bins of its tiles; wave-uniform.
__device__ __forceinline__ void bin_record(const PassDesc* passes, const Workspace& ws, bool have, uint4 r) {
    int tx0 = 0, ty0 = 0, nx = 0, n = 0, tile0 = 0, tiles_x = 0;
    if (have) {
        const PassDesc& P = passes[r.y];
        tx0 = (int)(r.z & 0xFFFFu);
        ty0 = (int)(r.z >> 16);
        nx = (int)(r.w & 0xFFFFu) - tx0 + 1;
        n = nx * ((int)(r.w >> 16) - ty0 + 1);
        tile0 = P.tile0;
        tiles_x = P.tiles_x;
    }
    uint64_t wide = __ballot(n > WIDE_TILES);
    while (wide) {
        int src = __ffsll((unsigned long long)wide) - 1;
        wide &= wide - 1;
        int wn = __shfl(n, src), wnx = __shfl(nx, src), wx0 = __shfl(tx0, src), wy0 = __shfl(ty0, src);
        int wt0 = __shfl(tile0, src), wtx = __shfl(tiles_x, src);
        uint32_t we = __shfl(r.x, src);
        for (int k = __lane_id(); k < wn; k += 64) {
            uint32_t pos = gadd(ws.tile_cursor + (wt0 + (wy0 + k / wnx) * wtx + wx0 + k % wnx), 1u);
            if (pos < ws.bin_cap) gst(ws.bins + pos, we);
        }
    }
    int ns = n > WIDE_TILES ? 0 : n;
    for (int k = 0; __ballot(k < ns); k++) {
        bool want = k < ns;
        int gt = want ? tile0 + (ty0 + k / nx) * tiles_x + tx0 + k % nx : 0;
        uint32_t pos = wave_reserve(ws.tile_cursor, gt, want);
        if (want && pos < ws.bin_cap) gst(ws.bins + pos, r.x);
    }
}

// Workgroups below nseg scatter the setup segment of their index: the
// segment's per-tile pair counts (listed by k_setup) go into an LDS table,
// each distinct tile reserves its range of the bin with one global atomic,
// then every pair takes a place in that range with an LDS atomic.  The other
// workgroups stride over the clipped fan triangles listed after the segments.
template <int BITS>
__global__ __launch_bounds__(BLOCK) void k_bin(const PassDesc* __restrict__ passes, Workspace ws, int nseg) {
    constexpr int BIN_HASH = 1 << BITS;
    if (ws.status[STATUS_SEQ_FLAGS] & FLAG_NO_RASTER) return;   // nothing is rasterised
    if ((int)blockIdx.x < nseg) {
        __shared__ uint32_t keys[BIN_HASH];
        __shared__ uint32_t cnt[BIN_HASH];   // a slot's next bin position, set when its tile is inserted
        // the segment's sizes (written by k_setup, uniform: scalar loads) are
        // read before the table is cleared, so that they are not awaited after it
        const uint32_t n = *(const HANA_CONST_AS uint32_t*)(ws.seg_count + blockIdx.x);
        const uint32_t info = *(const HANA_CONST_AS uint32_t*)(ws.seg_info + blockIdx.x);
        const uint4* rec = ws.tris + (size_t)blockIdx.x * (BLOCK * FPT);
        // a segment's triangles are all of one pass (one k_setup workgroup's faces)
        const uint32_t spass = n ? *(const HANA_CONST_AS uint32_t*)&rec[0].y : 0u;
        for (int i = threadIdx.x; i < BIN_HASH; i += BLOCK) keys[i] = BIN_EMPTY;
        __syncthreads();
        const PassDesc& P = passes[spass];
        // the segment's distinct tiles and their pair counts, as k_setup listed
        // them (a tile that found no slot there, or finds none here, goes
        // through the global cursor pair by pair below)
        uint32_t nlist = info & 0xFFFFu, nw = info >> 16;
        const uint2* list = ws.seg_tiles + (size_t)blockIdx.x * ws.seg_stride;
        // listed tiles are distinct, so an inserted slot is this thread's:
        // it reserves the tile's range of the bin at once
        for (uint32_t i = threadIdx.x; i < nlist; i += BLOCK) {
            uint2 e = gld2(list + i);
            int sl = bin_slot<BITS>(keys, e.x, true);
            if (sl >= 0) cnt[sl] = gadd(ws.tile_cursor + e.x, e.y);
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
            uint4 r = gld4(rec + i);
            int tx0 = (int)(r.z & 0xFFFFu), ty0 = (int)(r.z >> 16);
            int tx1 = (int)(r.w & 0xFFFFu), ty1 = (int)(r.w >> 16);
            if ((tx1 - tx0 + 1) * (ty1 - ty0 + 1) > WIDE_TILES) continue;
            for (int ty = ty0; ty <= ty1; ty++)
                for (int tx = tx0; tx <= tx1; tx++) {
                    uint32_t gt = (uint32_t)(P.tile0 + ty * P.tiles_x + tx);
                    int sl = bin_slot<BITS>(keys, gt, false);
                    uint32_t pos = sl >= 0 ? atomicAdd(&cnt[sl], 1u) : gadd(ws.tile_cursor + gt, 1u);
                    if (pos < ws.bin_cap) gst(ws.bins + pos, r.x);
                }
        }
        // wide triangles, one per wave at a time, their tiles spread over its lanes
        const uint16_t* wide = ws.seg_wide + (size_t)blockIdx.x * (BLOCK * FPT);
        for (uint32_t w = threadIdx.x >> 6; w < nw; w += BLOCK / 64) {
            uint4 r = gld4(rec + gld(wide + w));
            int tx0 = (int)(r.z & 0xFFFFu), ty0 = (int)(r.z >> 16);
            int nx = (int)(r.w & 0xFFFFu) - tx0 + 1;
            int nt = nx * ((int)(r.w >> 16) - ty0 + 1);
            for (int k = threadIdx.x & 63; k < nt; k += 64) {
                uint32_t pos = gadd(ws.tile_cursor + (P.tile0 + (ty0 + k / nx) * P.tiles_x + tx0 + k % nx), 1u);
                if (pos < ws.bin_cap) gst(ws.bins + pos, r.x);
            }
        }
        return;
    }
    uint32_t count = min(ws.status[STATUS_TRIS], ws.tris_cap - ws.seg_region);
    uint32_t stride = (gridDim.x - (uint32_t)nseg) * BLOCK;
    for (uint32_t base = (blockIdx.x - (uint32_t)nseg) * BLOCK; base < count; base += stride) {
        uint32_t i = base + threadIdx.x;
        bool have = i < count;
        bin_record(passes, ws, have, have ? gld4(ws.tris + (ws.seg_region + i)) : make_uint4(0, 0, 0, 0));
    }
}

// ---------------------------------------------------------------------------
// k_scan_reduce / k_scan_apply: exclusive scan of the tile counts into
// tile_offset (and tile_cursor, the scatter cursors).  Each workgroup owns
// SCAN_CHUNK counters; the apply kernel adds the sums of the chunks before.
// ---------------------------------------------------------------------------

// Sums are 64-bit so that a total past 2^32 pairs is seen, not wrapped.
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    uint32_t lo = __shfl_xor((uint32_t)v, m), hi = __shfl_xor((uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, int d) {
    uint32_t lo = __shfl_up((uint32_t)v, d), hi = __shfl_up((uint32_t)(v >> 32), d);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t block_sum(uint64_t v, uint64_t* sh) {
    for (int off = 32; off > 0; off >>= 1) v += shfl_xor64(v, off);
    int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    uint64_t t = 0;
    for (int i = 0; i < SCAN_THREADS / 64; i++) t += sh[i];
    return t;
}

// A thread's SCAN_PER_THREAD consecutive counters from b (a multiple of 16,
// so whole runs load as 16-byte vectors; buffers are 256-byte aligned).
__device__ __forceinline__ void load_counts(const uint32_t* c, int b, int n, uint32_t (&v)[SCAN_PER_THREAD]) {
    if (b + SCAN_PER_THREAD <= n) {
        const uint4* q = reinterpret_cast<const uint4*>(c + b);
#pragma unroll
        for (int i = 0; i < SCAN_PER_THREAD / 4; i++) {
            uint4 x = q[i];
            v[4 * i] = x.x; v[4 * i + 1] = x.y; v[4 * i + 2] = x.z; v[4 * i + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < SCAN_PER_THREAD; i++) v[i] = (b + i < n) ? c[b + i] : 0u;
    }
}

// Set-up and clipping are complete when the scan runs: their counts become
// the needs the host sizes the next launch sequence by.
__device__ __forceinline__ void note_clip_needs(const Workspace& ws) {
    atomicMax(&ws.status[STATUS_NEED_POLYS], ws.status[STATUS_POLYS]);
    atomicMax(&ws.status[STATUS_NEED_TRIS], ws.status[STATUS_TRIS]);
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(Workspace ws) {
    __shared__ uint64_t sh[SCAN_THREADS / 64];
    int n = ws.ntiles;
    int b = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER_THREAD;
    uint32_t v[SCAN_PER_THREAD];
    load_counts(ws.tile_count, b, n, v);
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER_THREAD; i++) s += v[i];
    uint64_t t = block_sum(s, sh);
    if (threadIdx.x == 0) ws.scan_part[blockIdx.x] = t;
    if (blockIdx.x == 0 && threadIdx.x == 0) note_clip_needs(ws);
}

// With a single chunk k_scan_reduce is not launched and this kernel alone
// scans the counts.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(Workspace ws) {
    __shared__ uint64_t sh[SCAN_THREADS / 64];
    __shared__ uint64_t wsum[SCAN_THREADS / 64];
    int n = ws.ntiles;
    int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (gridDim.x == 1 && tid == 0) note_clip_needs(ws);
    // sum of the chunks before this one
    uint64_t before = 0;
    for (int i = tid; i < (int)blockIdx.x; i += SCAN_THREADS) before += ws.scan_part[i];
    before = block_sum(before, sh);
    // this thread's counters, then an exclusive scan of the thread sums
    int b = blockIdx.x * SCAN_CHUNK + tid * SCAN_PER_THREAD;
    uint32_t v[SCAN_PER_THREAD];
    load_counts(ws.tile_count, b, n, v);
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER_THREAD; i++) s += v[i];
    uint64_t incl = s;
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t o = shfl_up64(incl, d);
        if (lane >= d) incl += o;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint64_t wbefore = 0;
    for (int i = 0; i < w; i++) wbefore += wsum[i];
    uint64_t run = before + wbefore + incl - s;
    // offsets past MAX_PAIRS are cut off (the sequence then rasterises nothing)
    if (b + SCAN_PER_THREAD <= n) {
        uint4* qo = reinterpret_cast<uint4*>(ws.tile_offset + b);
        uint4* qc = reinterpret_cast<uint4*>(ws.tile_cursor + b);
#pragma unroll
        for (int i = 0; i < SCAN_PER_THREAD / 4; i++) {
            uint4 x;
            x.x = (uint32_t)min(run, (uint64_t)MAX_PAIRS); run += v[4 * i];
            x.y = (uint32_t)min(run, (uint64_t)MAX_PAIRS); run += v[4 * i + 1];
            x.z = (uint32_t)min(run, (uint64_t)MAX_PAIRS); run += v[4 * i + 2];
            x.w = (uint32_t)min(run, (uint64_t)MAX_PAIRS); run += v[4 * i + 3];
            qo[i] = x;
            qc[i] = x;
        }
    } else {
        for (int i = 0; i < SCAN_PER_THREAD; i++)
            if (b + i < n) {
                ws.tile_offset[b + i] = (uint32_t)min(run, (uint64_t)MAX_PAIRS);
                ws.tile_cursor[b + i] = (uint32_t)min(run, (uint64_t)MAX_PAIRS);
                run += v[i];
            }
    }
    if (blockIdx.x == gridDim.x - 1 && tid == SCAN_THREADS - 1) {
        // `run` of the last thread is the grand total
        uint32_t total = (uint32_t)min(run, (uint64_t)MAX_PAIRS);
        ws.tile_offset[n] = total;
        ws.status[STATUS_BINS_NEEDED] = total;
        // past the pair limit no workspace would do: that flag alone, and no need
        // for the host to grow the bins by
        if (run > MAX_PAIRS) {
            raise_flag(ws, FLAG_PAIR_LIMIT);
        } else {
            atomicMax(&ws.status[STATUS_NEED_BINS], total);
            if (run > ws.bin_cap) raise_flag(ws, FLAG_BIN_OVERFLOW);
        }
        // every status word is final now (binning only reads them)
        if (ws.status_out)
            for (int i = 0; i < STATUS_WORDS; i++)
                ws.status_out[i] = __hip_atomic_load(&ws.status[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// k_raster: per-tile visibility in LDS, then one fragment program per pixel
// ---------------------------------------------------------------------------

constexpr int SMALL_PIX = HANA_SMALL_PIX;   // triangles with at most this many candidate pixels: one lane each

// k / w for 0 <= k < 1024 and 1 <= w <= 32 (a tile's pixels and row widths)
// as (k * magic) >> 16 with magic = ceil(2^16 / w): the excess
// k * (magic - 2^16/w) / 2^16 stays below 1024 / 2^16 = 1/64, less than the
// gap 1/w >= 1/32 between k/w and the next integer.
__device__ __forceinline__ uint32_t tile_div_magic(int w) { return (65536u + (uint32_t)w - 1u) / (uint32_t)w; }
__device__ __forceinline__ int tile_div(int k, uint32_t magic) { return (int)(((uint32_t)k * magic) >> 16); }

// The LDS of one raster workgroup: RB lanes (2 or 4 waves) over a tile of
// 2^TS pixels a side.
template <int TS, int RB>
struct TileLds {
    static constexpr int PIX = 1 << (2 * TS);
    static constexpr int LANES = RB;
    uint64_t key[PIX];
    uint32_t nan_order[PIX];   // 1 + order of the last NaN-depth fragment
    float ax[RB], ay[RB], s0x[RB], s0y[RB], s1x[RB], s1y[RB], uz[RB];
    float z0[RB], z1[RB], z2[RB];
    uint32_t lo[RB], box[RB];
    int big[RB], med[RB];
    // staging counters of sweep round r in set r % 3 (nbig, head, nmed, mhead):
    // round r clears set (r + 2) % 3 once every wave is past round r - 1
    int cnt[3][4];
    int nan_seen;
};

__device__ __forceinline__ bool load_tri(const PassDesc& P, const Workspace& ws, uint32_t e, TriS& t, uint32_t& order) {
    float4 a, b, c;
    if (!(e & CLIPPED_BIT)) {
        int f = (int)e;
        int i0 = gld(P.idx + 3 * f), i1 = gld(P.idx + 3 * f + 1), i2 = gld(P.idx + 3 * f + 2);
        a = gld4(P.vr + i0); b = gld4(P.vr + i1); c = gld4(P.vr + i2);
        order = (uint32_t)f << 4;
    } else {
        uint32_t pid = (e & ~CLIPPED_BIT) >> FAN_BITS, j = e & ((1u << FAN_BITS) - 1u);
        const float4* pv = ws.pool_vr + (size_t)pid * POLY_MAX;
        a = gld4(pv); b = gld4(pv + j + 1); c = gld4(pv + j + 2);
        order = ((uint32_t)gld(ws.pool_face + pid) << 4) + 1u + j;
    }
    return tri_setup(a, b, c, P.width, P.height, t);
}

// One fragment of triangle t at tile-local pixel (lx, ly).
template <int TS, int RB>
__device__ __forceinline__ void frag_resolve(TileLds<TS, RB>& L, const TriS& t, uint32_t lo, int xb, int yb, int lx, int ly,
                                             bool second) {
    float wx, wy, wz, z;
    if (!tri_cover_fast(t, xb + lx, yb + ly, wx, wy, wz, z)) return;
    int i = (ly << TS) | lx;
    if (!second) {
        if (z != z) {
            atomicMax(&L.nan_order[i], (0xFFFFFFFEu - lo) + 1u);
            L.nan_seen = 1;
            return;
        }
        atomicMin((unsigned long long*)&L.key[i], ((uint64_t)depth_bits(z) << 32) | lo);
    } else {
        uint32_t no = L.nan_order[i];
        if (no == 0u || z != z) return;
        uint32_t order = 0xFFFFFFFEu - lo;
        if (order + 1u <= no) return;   // only fragments after the last NaN one
        atomicMin((unsigned long long*)&L.key[i], ((uint64_t)depth_bits(z) << 32) | lo);
    }
}

// A triangle staged in LDS by lane s of the sweep, with its candidate box.
template <int TS, int RB>
__device__ __forceinline__ TriS staged_tri(const TileLds<TS, RB>& L, int s, uint32_t& lo, int& x0, int& y0, int& wdt,
                                           int& npix) {
    TriS t;
    t.ax = L.ax[s]; t.ay = L.ay[s]; t.s0x = L.s0x[s]; t.s0y = L.s0y[s];
    t.s1x = L.s1x[s]; t.s1y = L.s1y[s]; t.uz = L.uz[s];
    t.z0 = L.z0[s]; t.z1 = L.z1[s]; t.z2 = L.z2[s];
    lo = L.lo[s];
    uint32_t box = L.box[s];
    x0 = box & 0xFF;
    y0 = (box >> 16) & 0xFF;
    wdt = (int)((box >> 8) & 0xFF) - x0 + 1;
    npix = wdt * ((int)(box >> 24) - y0 + 1);
    return t;
}

// One sweep over the tile's bin: small triangles one per lane (in passes that
// share them, those of more than HANA_MED_PIX candidates HANA_MED_LANES lanes
// each), larger ones one per wave with the lanes spread over their candidate
// pixels.
template <int TS, int RB>
__device__ int sweep(TileLds<TS, RB>& L, const PassDesc& P, const Workspace& ws, uint32_t beg, uint32_t cnt, int xb, int yb,
                     bool second, int set) {
    constexpr int EDGE = 1 << TS;
    int tid = threadIdx.x, lane = tid & 63;
    for (uint32_t base = 0; base < cnt; base += RB) {
        int* C = L.cnt[set];
        bool ok = false;
        TriS t;
        uint32_t order = 0;
        if (base + tid < cnt) ok = load_tri(P, ws, gld(ws.bins + (beg + base + tid)), t, order);
        int bx0 = 0, bx1 = -1, by0 = 0, by1 = -1;
        if (ok) {
            bx0 = max(t.x0, xb) - xb; bx1 = min(t.x1, xb + EDGE - 1) - xb;
            by0 = max(t.y0, yb) - yb; by1 = min(t.y1, yb + EDGE - 1) - yb;
            ok = bx1 >= bx0 && by1 >= by0;
        }
        uint32_t lo = 0xFFFFFFFEu - order;
        if (ok) {
            int wdt = bx1 - bx0 + 1;
            int npix = wdt * (by1 - by0 + 1);
            if (npix <= SMALL_PIX && (!P.share_small || npix <= HANA_MED_PIX)) {
                for (int ly = by0; ly <= by1; ly++)
                    for (int lx = bx0; lx <= bx1; lx++) frag_resolve<TS, RB>(L, t, lo, xb, yb, lx, ly, second);
            } else {
                L.ax[tid] = t.ax; L.ay[tid] = t.ay; L.s0x[tid] = t.s0x; L.s0y[tid] = t.s0y;
                L.s1x[tid] = t.s1x; L.s1y[tid] = t.s1y; L.uz[tid] = t.uz;
                L.z0[tid] = t.z0; L.z1[tid] = t.z1; L.z2[tid] = t.z2;
                L.lo[tid] = lo;
                L.box[tid] = (uint32_t)bx0 | ((uint32_t)bx1 << 8) | ((uint32_t)by0 << 16) | ((uint32_t)by1 << 24);
                if (npix <= SMALL_PIX) L.med[atomicAdd(&C[2], 1)] = tid;
                else L.big[atomicAdd(&C[0], 1)] = tid;
            }
        }
        __syncthreads();
        // every wave is past round r - 1: its set is cleared for round r + 2
        const int nmed = C[2], nbig = C[0];
        if (tid == 0) {
            int* R = L.cnt[set == 0 ? 2 : set - 1];
            R[0] = 0; R[1] = 0; R[2] = 0; R[3] = 0;
        }
        // shared small triangles: 64 / MED_LANES per wave and round
        for (;;) {
            constexpr int PER_WAVE = 64 / HANA_MED_LANES;
            if (nmed == 0) break;
            int q = 0;
            if (lane == 0) q = atomicAdd(&C[3], PER_WAVE);
            q = __shfl(q, 0);
            if (q >= nmed) break;
            int m = q + lane / HANA_MED_LANES;
            if (m < nmed) {
                uint32_t mlo;
                int x0, y0, wdt, npix;
                TriS mt = staged_tri<TS, RB>(L, L.med[m], mlo, x0, y0, wdt, npix);
                uint32_t magic = tile_div_magic(wdt);
                for (int k = lane % HANA_MED_LANES; k < npix; k += HANA_MED_LANES) {
                    int row = tile_div(k, magic);
                    frag_resolve<TS, RB>(L, mt, mlo, xb, yb, x0 + k - row * wdt, y0 + row, second);
                }
            }
        }
        for (;;) {
            if (nbig == 0) break;
            int q = 0;
            if (lane == 0) q = atomicAdd(&C[1], 1);
            q = __shfl(q, 0);
            if (q >= nbig) break;
            uint32_t blo;
            int x0, y0, wdt, npix;
            TriS bt = staged_tri<TS, RB>(L, L.big[q], blo, x0, y0, wdt, npix);
            uint32_t magic = tile_div_magic(wdt);
            for (int k = lane; k < npix; k += 64) {
                int row = tile_div(k, magic);
                frag_resolve<TS, RB>(L, bt, blo, xb, yb, x0 + k - row * wdt, y0 + row, second);
            }
        }
        // staged triangles are read until every wave is done with them (a
        // round that staged none wrote all its fragments before the barrier above)
        if (nmed | nbig) __syncthreads();
        set = set == 2 ? 0 : set + 1;
    }
    return set;
}

// Eight varyings of one corner (a 32-byte aligned row) in two 16-byte loads.
__device__ __forceinline__ void load_varyings8(const float* row, float* y) {
    float4 lo = gld4(reinterpret_cast<const float4*>(row)), hi = gld4(reinterpret_cast<const float4*>(row) + 1);
    y[0] = lo.x; y[1] = lo.y; y[2] = lo.z; y[3] = lo.w;
    y[4] = hi.x; y[5] = hi.y; y[6] = hi.z; y[7] = hi.w;
}

// Interpolates, shades and writes the winning fragment of one pixel
// (interpolate_varyings + fragment + set_depth/set_color, SR/graphics.cpp:205-220, 362-373).
__device__ __forceinline__ void shade_pixel(const PassDesc& P, const Workspace& ws, uint32_t order, int px, int py) {
    uint32_t f = order >> 4, slot = order & 15u;
    float4 va, vb, vc;
    const float *sa, *sb, *sc;
    int nv;
    if (slot == 0) {
        int i0 = gld(P.idx + 3 * f), i1 = gld(P.idx + 3 * f + 1), i2 = gld(P.idx + 3 * f + 2);
        va = gld4(P.vr + i0); vb = gld4(P.vr + i1); vc = gld4(P.vr + i2);
        nv = P.nvary;
        sa = P.vs + (size_t)i0 * nv; sb = P.vs + (size_t)i1 * nv; sc = P.vs + (size_t)i2 * nv;
    } else {
        int pid = gld(P.poly_of_face + f);
        uint32_t j = slot - 1;
        size_t base = (size_t)pid * POLY_MAX;
        va = gld4(ws.pool_vr + base); vb = gld4(ws.pool_vr + base + j + 1); vc = gld4(ws.pool_vr + base + j + 2);
        nv = POLY_VARY;
        sa = ws.pool_vs + base * POLY_VARY;
        sb = ws.pool_vs + (base + j + 1) * POLY_VARY;
        sc = ws.pool_vs + (base + j + 2) * POLY_VARY;
    }
    // the corners' live varyings, all loaded before any is awaited (8 per
    // corner as two 16-byte loads: their rows are 32-byte aligned)
    int cnt = varyings_of(P.kind);
    float ya[8], yb[8], yc[8];
    if (cnt == 8) {
        load_varyings8(sa, ya);
        load_varyings8(sb, yb);
        load_varyings8(sc, yc);
    } else {
        for (int c = 0; c < 8; c++) ya[c] = yb[c] = yc[c] = 0.0f;
        ya[0] = gld(sa); yb[0] = gld(sb); yc[0] = gld(sc);
        if (cnt == 2) { ya[1] = gld(sa + 1); yb[1] = gld(sb + 1); yc[1] = gld(sc + 1); }
    }
    TriS t;
    tri_setup(va, vb, vc, P.width, P.height, t);
    // The weights take the same quotients as the resolve.  Where a numerator
    // is zero, tri_cover_fast may give a zero weight of the other sign than the
    // division: wx = 1 - q is then 1 either way, and a zero wy or wz only adds
    // a signed zero beside the nonzero va.w * wx in the normaliser and in each
    // varying's sum, so no sum changes unless it is exactly zero, and then only
    // in the sign of that zero.  Depth is summed from +0 and does not change.
    float wx, wy, wz, z;
    tri_cover_fast(t, px, py, wx, wy, wz, z);
    float k0 = va.w * wx, k1 = vb.w * wy, k2 = vc.w * wz;
    float normalizer = 1 / (k0 + k1 + k2);
    float vy[8];
    for (int c = 0; c < 8; c++) {
        if (c < cnt) {
            float sum = ya[c] * k0 + yb[c] * k1 + yc[c] * k2;
            vy[c] = sum * normalizer;
        } else {
            vy[c] = 0.0f;
        }
    }
    Tex dt{P.diffuse, P.dw, P.dh}, nt{P.normal, P.nw, P.nh};
    ShadowMap sm{P.shadow, P.shadow_depth, P.sw, P.sh};
    size_t pix = (size_t)py * P.width + px;
    // the fp32 extension's shadow pass is depth-only
    if (!(P.kind == SHADER_SHADOW && uniforms_of(P.u).shadow_fp32)) {
        C3 col = fragment_program(P.kind, uniforms_of(P.u), vy, dt, nt, sm);
        uint32_t alpha = P.clear ? P.clear_rgba : gld(P.color + pix);
        gst(P.color + pix, pack_rgb(col, alpha));
    } else if (P.clear) {
        gst(P.color + pix, P.clear_rgba);
    }
    // A NaN depth is stored as the x86 default NaN (quiet, sign set): that is
    // what the reference's own arithmetic produces for it (an invalid
    // operation such as inf/inf yields that pattern and NaN operands pass it on).
    gst(P.depth + pix, (z != z) ? __uint_as_float(0xFFC00000u) : z);
    if (P.winner) gst(P.winner + pix, order);
}

// Workgroups are dispatched to the 8 XCDs round-robin by blockIdx.  The
// logical tile of a workgroup is chosen so that each XCD receives whole runs
// of HANA_XCD_CHUNK consecutive tiles (a piece of a tile row), the runs dealt
// to the XCDs in turn: vertices, bins and texels of a run stay in one XCD's L2
// while every XCD still gets tiles from all over the target (objects cover the
// middle of a frame more than its edges).  Chunk 0 keeps blockIdx order.
#ifndef HANA_XCD_CHUNK
#define HANA_XCD_CHUNK 4
#endif
__device__ __forceinline__ int xcd_logical_block(int b, int nb) {
    const int kXcd = 8, ch = HANA_XCD_CHUNK;
    if (ch <= 0) return b;
    int full = nb / (kXcd * ch) * (kXcd * ch);
    if (b >= full) return b;
    int xcd = b % kXcd, slot = b / kXcd;
    return ((slot / ch) * kXcd + xcd) * ch + slot % ch;
}

// Renders tile lt of pass P: the visibility of its pixels in LDS, then one
// fragment program per covered pixel; a tile no triangle reaches only takes
// the pass's clear.  The workgroup's threads all call it; it leaves L in use
// until they have all returned.
template <int TS, int RB>
__device__ void raster_tile(TileLds<TS, RB>& L, const PassDesc& P, const Workspace& ws, int lt, uint32_t beg,
                            uint32_t cnt) {
    constexpr int EDGE = 1 << TS, PIX = 1 << (2 * TS);
    int tx = lt % P.tiles_x, ty = lt / P.tiles_x;
    int xb = tx << TS, yb = ty << TS;
    int tid = threadIdx.x;
    int W = P.width, H = P.height;

    if (beg >= ws.bin_cap) cnt = 0;
    else if (cnt > ws.bin_cap - beg) cnt = ws.bin_cap - beg;
    if (cnt == 0) {   // uniform: no fragment, the clear only
        for (int i = tid; i < PIX; i += RB) {
            int px = xb + (i & (EDGE - 1)), py = yb + (i >> TS);
            if (px >= W || py >= H) continue;
            size_t pix = (size_t)py * W + px;
            if (P.clear) {
                gst(P.color + pix, P.clear_rgba);
                gst(P.depth + pix, P.clear_depth);
            }
            if (P.winner) gst(P.winner + pix, NO_WINNER);
        }
        return;
    }

    for (int i = tid; i < PIX; i += RB) {
        int px = xb + (i & (EDGE - 1)), py = yb + (i >> TS);
        uint64_t k = ~0ull;
        if (px < W && py < H) k = init_key(P.clear ? P.clear_depth : gld(P.depth + ((size_t)py * W + px)));
        L.key[i] = k;
        L.nan_order[i] = 0u;
    }
    if (tid < 12) (&L.cnt[0][0])[tid] = 0;
    if (tid == 0) L.nan_seen = 0;
    __syncthreads();

    int set = sweep<TS, RB>(L, P, ws, beg, cnt, xb, yb, false, 0);

    if (L.nan_seen) {
        // pixels that saw a NaN-depth fragment restart after the last one
        for (int i = tid; i < PIX; i += RB) {
            uint32_t no = L.nan_order[i];
            if (no) L.key[i] = (0xFFFFFFFFull << 32) | (uint64_t)(0xFFFFFFFEu - (no - 1u));
        }
        __syncthreads();
        sweep<TS, RB>(L, P, ws, beg, cnt, xb, yb, true, set);
    }

    uint32_t written = 0;
    for (int i = tid; i < PIX; i += RB) {
        int px = xb + (i & (EDGE - 1)), py = yb + (i >> TS);
        if (px >= W || py >= H) continue;
        uint32_t lo = (uint32_t)L.key[i];
        size_t pix = (size_t)py * W + px;
        if (lo == 0xFFFFFFFFu) {
            if (P.clear) {
                gst(P.color + pix, P.clear_rgba);
                gst(P.depth + pix, P.clear_depth);
            }
            if (P.winner) gst(P.winner + pix, NO_WINNER);
            continue;
        }
        shade_pixel(P, ws, 0xFFFFFFFEu - lo, px, py);
        written++;
    }
    if (P.stats) stat_add(P, ST_PIXELS, written);
}

// The bin range of global tile gt (uniform; written by earlier launches of
// the sequence only, so read through the scalar cache).
__device__ __forceinline__ uint2 tile_range(const Workspace& ws, int gt) {
    return make_uint2(*(const HANA_CONST_AS uint32_t*)(ws.tile_offset + gt),
                      *(const HANA_CONST_AS uint32_t*)(ws.tile_count + gt));
}

// Waves per SIMD that the raster's LDS allows (160 KiB per CU, 4 SIMDs, at
// most 8 waves each).  The kernels ask the compiler for that occupancy: the
// persistent raster, whose tile loop keeps more values live, would otherwise
// take 78-84 VGPRs (6 waves per SIMD) where its LDS allows 8.
template <int TS, int RB>
constexpr int raster_occupancy() {
    constexpr int lds = (int)sizeof(TileLds<TS, RB>) + 16;
    constexpr int w = (163840 / lds) * (RB / 64) / 4;
    return w < 1 ? 1 : (w > 8 ? 8 : w);
}

// The pass of raster workgroup blk of group GROUP and its global tile: one
// scalar load of the group's lookup (Workspace::tile_pass), then, in a run of
// workgroups that spans passes, a step per pass boundary; a binary search
// without a lookup.
template <int GROUP>
__device__ __forceinline__ int raster_pass(const PassDesc* P, int n, const Workspace& ws, int blk, int& gt) {
    const uint2* tp = ws.tile_pass[GROUP];
    if (!tp) {
        int p = find_pass<2>(P, n, blk);
        gt = P[p].tile0 + blk - P[p].tblock0;
        return p;
    }
    u2_vec e = *(const HANA_CONST_AS u2_vec*)(tp + (blk >> TILE_PASS_SHIFT));
    int p = (int)(e.x & ~TILE_PASS_PURE);
    gt = blk + (int)e.y;
    if (!(e.x & TILE_PASS_PURE)) {
        while (p + 1 < n && P[p + 1].tblock0 <= blk) p++;
        gt = P[p].tile0 + blk - P[p].tblock0;
    }
    return p;
}

// One workgroup per tile of one dependency group (passes of the group, their
// tiles numbered from 0 by tblock0), XCD-banded.  GROUP only names the launch
// (the shadow passes' group 0, the main passes' group 1) in profiles.
template <int TS, int GROUP, int RB>
__global__ __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(raster_occupancy<TS, RB>())))
void k_raster(const PassDesc* __restrict__ passes, int npasses, Workspace ws) {
    __shared__ TileLds<TS, RB> L;
    // the loads that decide the workgroup's work are issued before any of
    // them is awaited: the sequence flags, the pass lookup, the bin range
    uint32_t flags = *(const HANA_CONST_AS uint32_t*)(ws.status + STATUS_SEQ_FLAGS);
    int blk = xcd_logical_block(blockIdx.x, gridDim.x);
    int gt;
    const PassDesc& P = passes[raster_pass<GROUP>(passes, npasses, ws, blk, gt)];
    uint2 range = tile_range(ws, gt);
    // an incomplete sequence leaves its targets untouched: the host renders it
    // again.  (Bin offsets never reach 2^32 - 1; testing the range as well
    // keeps its loads ahead of the flags' wait.)
    if ((flags & FLAG_NO_RASTER) | (uint32_t)((range.x & range.y) == 0xFFFFFFFFu)) return;
    raster_tile<TS, RB>(L, P, ws, blk - P.tblock0, range.x, range.y);
}

// The persistent raster: a fixed grid of workgroups takes the tiles of every
// pass of the sequence from one queue, RQ_CHUNK consecutive tiles at a time
// (they share triangles, which stay in the taking XCD's L2).  Queue items are
// global tile indices: the shadow passes' tiles first, then the main passes
// that sample no shadow map drawn here, then those that do.  A main pass of
// the last kind needs every shadow tile finished:
//  - a workgroup counts the shadow tiles it finishes and publishes the count
//    once, with one release at agent scope (its shadow-map writes become
//    visible on every XCD), when it first takes an item past them or finds
//    the queue empty;
//  - before its first item that samples a map, it waits until the published
//    total is every shadow tile, then acquires once.
// Items are taken in queue order, so every shadow tile was taken before any
// item that waits, by a workgroup that publishes before it could wait itself:
// the wait ends.  It is still bounded: past RQ_SPIN_LIMIT polls it raises
// FLAG_QUEUE_STALL (an error for the host) and goes on.
constexpr uint32_t RQ_CHUNK = 4;
constexpr uint32_t RQ_SPIN_LIMIT = 1u << 24;

// largest pass index whose first global tile is <= item (passes sorted by it)
__device__ __forceinline__ int pass_of_item(const PassDesc* P, int n, uint32_t item) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if ((uint32_t)P[mid].tile0 <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int TS, int RB>
__global__ __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(raster_occupancy<TS, RB>())))
void k_raster_q(const PassDesc* __restrict__ passes, int npasses, Workspace ws) {
    __shared__ TileLds<TS, RB> L;
    __shared__ uint32_t s_first;
    if (ws.status[STATUS_SEQ_FLAGS] & FLAG_NO_RASTER) return;
    const int tid = threadIdx.x;
    uint32_t mine = 0;   // shadow tiles finished here, not yet published
    bool published = false, ready = false;
    for (;;) {
        if (tid == 0) s_first = atomicAdd(&ws.qstate[0], RQ_CHUNK);
        __syncthreads();
        // uniform: the tile's pass, bounds and tables then live in SGPRs
        uint32_t first = __builtin_amdgcn_readfirstlane(s_first);
        __syncthreads();
        uint32_t end = min(first + RQ_CHUNK, ws.q_items);
        if (!published && (first >= ws.q_g0 || first >= ws.q_items)) {
            if (tid == 0 && mine)
                __hip_atomic_fetch_add(&ws.qstate[1], mine, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            published = true;
        }
        if (first >= ws.q_items) break;
        for (uint32_t it = first; it < end; it++) {
            if (!published && it >= ws.q_g0) {
                if (tid == 0 && mine)
                    __hip_atomic_fetch_add(&ws.qstate[1], mine, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                published = true;
            }
            if (!ready && it >= ws.q_dep_from) {
                if (tid == 0) {
                    uint32_t spins = 0;
                    while (__hip_atomic_load(&ws.qstate[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < ws.q_g0) {
                        if (++spins > RQ_SPIN_LIMIT) {
                            raise_flag(ws, FLAG_QUEUE_STALL);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(8);
                    }
                }
                __syncthreads();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                ready = true;
            }
            const PassDesc& P = passes[pass_of_item(passes, npasses, it)];
            uint2 range = tile_range(ws, (int)it);
            raster_tile<TS, RB>(L, P, ws, (int)(it - (uint32_t)P.tile0), range.x, range.y);
            __syncthreads();   // the tile's writes are issued and L is free again
            if (it < ws.q_g0) mine++;
        }
    }
}

// fills a 32-bit buffer (graph-safe clear)
__global__ void k_fill32(uint32_t* dst, uint32_t value, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = value;
}

// window_draw_buffer (SR/win32.cpp:348-370) without the window: row r of the
// target lands on row h-1-r, R and B bytes swap (RGBA -> BGRA), alpha is
// carried along.  One word per lane, coalesced reads and writes per row.
__global__ void k_present(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, int w, int h) {
    size_t n = (size_t)w * h;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        size_t r = i / (size_t)w, c = i - r * (size_t)w;
        uint32_t v = src[i];
        dst[((size_t)(h - 1) - r) * (size_t)w + c] = ((v & 0xFFu) << 16) | (v & 0xFF00FF00u) | ((v >> 16) & 0xFFu);
    }
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------

hipError_t launch_vertex(const PassDesc* d_passes, int npasses, const Workspace& ws, int blocks, hipStream_t s) {
    if (npasses <= 0) return hipSuccess;
    // k_vertex also zeroes the tile counters and the status words of the
    // sequence, so it runs even when no pass has a vertex
    blocks = std::max(blocks, 1);
    hipLaunchKernelGGL(k_vertex, dim3(blocks), dim3(BLOCK), 0, s, d_passes, npasses, ws);
    return hipGetLastError();
}
hipError_t launch_setup(const PassDesc* d_passes, int npasses, const Workspace& ws, int blocks, bool small_tris,
                        hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    if (ws.tile_shift == 5) {
        if (small_tris) hipLaunchKernelGGL((k_setup<9, 5>), dim3(blocks), dim3(BLOCK), 0, s, d_passes, npasses, ws);
        else hipLaunchKernelGGL((k_setup<10, 5>), dim3(blocks), dim3(BLOCK), 0, s, d_passes, npasses, ws);
    } else {
        if (small_tris) hipLaunchKernelGGL((k_setup<9, 4>), dim3(blocks), dim3(BLOCK), 0, s, d_passes, npasses, ws);
        else hipLaunchKernelGGL((k_setup<10, 4>), dim3(blocks), dim3(BLOCK), 0, s, d_passes, npasses, ws);
    }
    return hipGetLastError();
}
hipError_t launch_clip(const PassDesc* d_passes, const Workspace& ws, hipStream_t s) {
    // grid-stride: the queue length is only known on the device
    if (ws.tile_shift == 5) hipLaunchKernelGGL(k_clip<5>, dim3(2048), dim3(CLIP_THREADS), 0, s, d_passes, ws);
    else hipLaunchKernelGGL(k_clip<4>, dim3(2048), dim3(CLIP_THREADS), 0, s, d_passes, ws);
    return hipGetLastError();
}
int scan_chunks(int ntiles) { return ntiles > 0 ? (ntiles + SCAN_CHUNK - 1) / SCAN_CHUNK : 1; }
hipError_t launch_scan(const Workspace& ws, hipStream_t s) {
    int chunks = scan_chunks(ws.ntiles);
    if (chunks > 1) {
        hipLaunchKernelGGL(k_scan_reduce, dim3(chunks), dim3(SCAN_THREADS), 0, s, ws);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_scan_apply, dim3(chunks), dim3(SCAN_THREADS), 0, s, ws);
    return hipGetLastError();
}
hipError_t launch_bin(const PassDesc* d_passes, const Workspace& ws, int nseg, bool small_tris, hipStream_t s) {
    // one workgroup per setup segment, then a grid-stride tail for the
    // clipped fan triangles (their count is only known on the device)
    if (small_tris)
        hipLaunchKernelGGL(k_bin<9>, dim3(nseg + 512), dim3(BLOCK), 0, s, d_passes, ws, nseg);
    else
        hipLaunchKernelGGL(k_bin<11>, dim3(nseg + 512), dim3(BLOCK), 0, s, d_passes, ws, nseg);
    return hipGetLastError();
}
template <int TS, int RB>
static void raster_launch(const PassDesc* d_passes, int npasses, const Workspace& ws, int blocks, int group,
                          hipStream_t s) {
    if (group) hipLaunchKernelGGL((k_raster<TS, 1, RB>), dim3(blocks), dim3(RB), 0, s, d_passes, npasses, ws);
    else hipLaunchKernelGGL((k_raster<TS, 0, RB>), dim3(blocks), dim3(RB), 0, s, d_passes, npasses, ws);
}
hipError_t launch_raster(const PassDesc* d_passes, int npasses, const Workspace& ws, int blocks, int group,
                         hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    bool two = ws.raster_waves == 2;
    if (ws.tile_shift == 5) {
        if (two) raster_launch<5, 128>(d_passes, npasses, ws, blocks, group, s);
        else raster_launch<5, 256>(d_passes, npasses, ws, blocks, group, s);
    } else {
        if (two) raster_launch<4, 128>(d_passes, npasses, ws, blocks, group, s);
        else raster_launch<4, 256>(d_passes, npasses, ws, blocks, group, s);
    }
    return hipGetLastError();
}
hipError_t launch_raster_queue(const PassDesc* d_passes, int npasses, const Workspace& ws, int workgroups,
                               hipStream_t s) {
    if (workgroups <= 0 || ws.q_items == 0) return hipSuccess;
    bool two = ws.raster_waves == 2;
    if (ws.tile_shift == 5) {
        if (two) hipLaunchKernelGGL((k_raster_q<5, 128>), dim3(workgroups), dim3(128), 0, s, d_passes, npasses, ws);
        else hipLaunchKernelGGL((k_raster_q<5, 256>), dim3(workgroups), dim3(256), 0, s, d_passes, npasses, ws);
    } else {
        if (two) hipLaunchKernelGGL((k_raster_q<4, 128>), dim3(workgroups), dim3(128), 0, s, d_passes, npasses, ws);
        else hipLaunchKernelGGL((k_raster_q<4, 256>), dim3(workgroups), dim3(256), 0, s, d_passes, npasses, ws);
    }
    return hipGetLastError();
}
hipError_t launch_present(const uint32_t* src, uint32_t* dst, int w, int h, hipStream_t s) {
    size_t n = (size_t)w * h;
    if (n == 0) return hipSuccess;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_present, dim3(blocks), dim3(256), 0, s, src, dst, w, h);
    return hipGetLastError();
}

hipError_t launch_fill32(uint32_t* dst, uint32_t value, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fill32, dim3(blocks), dim3(256), 0, s, dst, value, n);
    return hipGetLastError();
}

}  // namespace hana
