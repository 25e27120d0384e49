// hana_device.h -- device-side arithmetic and shader programs for gfx950.
//
// Every expression keeps the reference's fp32 association order and the
// library is compiled with -ffp-contract=off, so products and sums round
// exactly as the reference's do (SR/x:N = Hana-SoftwareRenderer/x line N):
//   dot products accumulate in reverse index order from +0     SR/vector.h:69-73
//   norms accumulate forward                                    SR/vector.h:41-42
//   Color saturating arithmetic with std::min/max NaN rules     SR/color.cpp:38-63
// Division and sqrt are the correctly rounded IEEE operations
// (-fhip-fp32-correctly-rounded-divide-sqrt).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hana_types.h"

namespace hana {

// Address spaces of the loads.  Pointers held in descriptor tables are
// generic, so plain loads through them compile to FLAT instructions (their
// completion is counted with the LDS operations'), and loads of uniform data
// that follow a store in the same kernel must be repeated per lane.
//  - Uniform blocks never change during a launch: they are read through the
//    constant address space, so the compiler keeps them in SGPRs.
//  - Per-lane data (meshes, post-transform vertices, bins, textures, shadow
//    maps) is read through the global address space.
#define HANA_CONST_AS __attribute__((address_space(4)))
#define HANA_GLOBAL_AS __attribute__((address_space(1)))
typedef const HANA_CONST_AS DevUniforms& Uniforms;
typedef const HANA_CONST_AS float* UniformRow;
__device__ __forceinline__ Uniforms uniforms_of(const DevUniforms* p) { return *(const HANA_CONST_AS DevUniforms*)p; }

template <class T>
__device__ __forceinline__ T gld(const T* p) { return *(const HANA_GLOBAL_AS T*)p; }
template <class T>
__device__ __forceinline__ void gst(T* p, T v) { *(HANA_GLOBAL_AS T*)p = v; }
typedef float f4_vec __attribute__((ext_vector_type(4)));
typedef float f2_vec __attribute__((ext_vector_type(2)));
typedef uint32_t u4_vec __attribute__((ext_vector_type(4)));
typedef uint32_t u2_vec __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float4 gld4(const float4* p) {
    f4_vec v = *(const HANA_GLOBAL_AS f4_vec*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 gld4(const uint4* p) {
    u4_vec v = *(const HANA_GLOBAL_AS u4_vec*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 gld2(const uint2* p) {
    u2_vec v = *(const HANA_GLOBAL_AS u2_vec*)p;
    return make_uint2(v.x, v.y);
}
__device__ __forceinline__ void gst4(float4* p, float4 v) {
    f4_vec w = {v.x, v.y, v.z, v.w};
    *(HANA_GLOBAL_AS f4_vec*)p = w;
}
__device__ __forceinline__ void gst4(uint4* p, uint4 v) {
    u4_vec w = {v.x, v.y, v.z, v.w};
    *(HANA_GLOBAL_AS u4_vec*)p = w;
}
__device__ __forceinline__ void gst2(float2* p, float2 v) {
    f2_vec w = {v.x, v.y};
    *(HANA_GLOBAL_AS f2_vec*)p = w;
}
__device__ __forceinline__ void gst2(uint2* p, uint2 v) {
    u2_vec w = {v.x, v.y};
    *(HANA_GLOBAL_AS u2_vec*)p = w;
}
// atomicAdd at agent scope, relaxed, as a global (not FLAT) atomic
__device__ __forceinline__ uint32_t gadd(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add((HANA_GLOBAL_AS uint32_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct F3 { float x, y, z; };

__device__ __forceinline__ F3 f3(float x, float y, float z) { F3 v; v.x = x; v.y = y; v.z = z; return v; }
__device__ __forceinline__ F3 sub3(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 add3(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 mul3(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }

// operator*(Vector,Vector) (SR/vector.h:69-73)
__device__ __forceinline__ float dot3(F3 a, F3 b) {
    float r = 0.0f;
    r += a.z * b.z;
    r += a.y * b.y;
    r += a.x * b.x;
    return r;
}
// Matrix row * (x, y, z, w) (SR/matrix.h:112-116); the row may live in the
// constant address space (uniform blocks)
template <class Row>
__device__ __forceinline__ float dot4(Row row, float x, float y, float z, float w) {
    float r = 0.0f;
    r += row[3] * w;
    r += row[2] * z;
    r += row[1] * y;
    r += row[0] * x;
    return r;
}
// Vector<3>::normalize (SR/vector.h:41-42)
__device__ __forceinline__ F3 normalize3(F3 v) {
    float n = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    float s = 1.0f / n;
    return mul3(v, s);
}
// cross (SR/vector.h:97-99)
__device__ __forceinline__ F3 cross3(F3 a, F3 b) {
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }   // std::max
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }   // std::min
__device__ __forceinline__ float saturate(float f) { return f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f); }  // SR/maths.cpp:7-9

// float -> int with the x86-64 result the reference gets (truncation,
// INT_MIN for NaN and out-of-range values).
__device__ __forceinline__ int f2i(float f) {
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return (int)0x80000000u;
    return (int)f;
}

// ---- Color (SR/color.cpp:38-63); alpha is never stored, so it is dropped ----
struct C3 { float r, g, b; };
__device__ __forceinline__ C3 c3(float r, float g, float b) { C3 c; c.r = r; c.g = g; c.b = b; return c; }
__device__ __forceinline__ float clamp01s(float v) { return smin(smax(0.0f, v), 1.0f); }
__device__ __forceinline__ C3 cadd(C3 a, C3 b) { return c3(clamp01s(a.r + b.r), clamp01s(a.g + b.g), clamp01s(a.b + b.b)); }
__device__ __forceinline__ C3 cmulf(C3 a, float i) {
    i = (i > 1.0f ? 1.0f : (i < 0.0f ? 0.0f : i));
    return c3(clamp01s(a.r * i), clamp01s(a.g * i), clamp01s(a.b * i));
}
__device__ __forceinline__ C3 cmulc(C3 a, C3 b) { return c3(a.r * b.r, a.g * b.g, a.b * b.b); }
template <class Ptr>
__device__ __forceinline__ C3 c3p(Ptr p) { return c3(p[0], p[1], p[2]); }

// b / 255.f for a byte b, the correctly rounded quotient: the product with
// RN(1/255) and one fma correction, which equal the IEEE division for every
// one of the 256 bytes (checked exhaustively on the host), in 3 instructions
// where the division takes 11.
__device__ __forceinline__ float byte_div255(uint32_t b) {
    const float r = 0x1.010102p-8f;   // RN(1/255)
    float x = (float)b;
    float q = x * r;
    float e = fmaf(-255.0f, q, x);
    return fmaf(e, r, q);
}

// ---- textures: RGBA8 words holding TGAColor bytes b,g,r,a (SR/tgaimage.h:46-53) ----
struct Tex {
    const uint32_t* texels;
    int w, h;
};
// TGAImage::get: out of range or no data -> all-zero TGAColor (SR/tgaimage.cpp:248-253)
__device__ __forceinline__ uint32_t tex_fetch(const Tex& t, float u, float v) {
    int x = f2i(u * (float)t.w), y = f2i(v * (float)t.h);
    if (!t.texels || x < 0 || y < 0 || x >= t.w || y >= t.h) return 0u;
    return gld(t.texels + ((size_t)y * (size_t)t.w + (size_t)x));
}
// Color(TGAColor) (SR/color.cpp:5) for IShader::tex_diffuse (SR/IShader.h:85-89)
__device__ __forceinline__ C3 tex_diffuse(const Tex& t, float u, float v) {
    uint32_t c = tex_fetch(t, u, v);
    return c3(byte_div255((c >> 16) & 0xFF), byte_div255((c >> 8) & 0xFF), byte_div255(c & 0xFF));
}
// IShader::tex_normal: res[2-i] = c[i]/255*2-1 (SR/IShader.h:91-99)
__device__ __forceinline__ F3 tex_normal(const Tex& t, float u, float v) {
    uint32_t c = tex_fetch(t, u, v);
    F3 r;
    r.z = byte_div255(c & 0xFF) * 2.f - 1.f;
    r.y = byte_div255((c >> 8) & 0xFF) * 2.f - 1.f;
    r.x = byte_div255((c >> 16) & 0xFF) * 2.f - 1.f;
    return r;
}

// number of interpolated floats a shader kind consumes (its live v2f fields)
__host__ __device__ __forceinline__ int varyings_of(int kind) {
    switch (kind) {
    case SHADER_GROUND: case SHADER_TOON: case SHADER_SHADOW: return 1;
    case SHADER_TEXTURE: return 2;
    default: return 8;
    }
}

// ---- vertex programs (SR/IShader.cpp:5-174) ----
// clip: clip position; vary: the live v2f fields of the kind, in the order
//   BLINN/NORMALMAP : world_pos.xyz, world_normal.xyz, uv
//   TEXTURE_LIGHT   : world_normal.xyz, uv
//   TEXTURE         : uv
//   GROUND/TOON     : intensity
//   SHADOW          : clip_pos.z
__device__ __forceinline__ void vertex_program(int kind, Uniforms u, F3 p, F3 n_obj, float tu, float tv,
                                               float clip[4], float vary[8]) {
    UniformRow m = (kind == SHADER_SHADOW) ? u.light_m : u.clip_m;   // ObjectToViewPos / ObjectToClipPos
    clip[3] = dot4(m + 12, p.x, p.y, p.z, 1.0f);
    clip[2] = dot4(m + 8, p.x, p.y, p.z, 1.0f);
    clip[1] = dot4(m + 4, p.x, p.y, p.z, 1.0f);
    clip[0] = dot4(m + 0, p.x, p.y, p.z, 1.0f);
    if (kind == SHADER_SHADOW) { vary[0] = clip[2]; return; }
    if (kind == SHADER_TEXTURE) { vary[0] = tu; vary[1] = tv; return; }
    // Model::normal renormalises (SR/model.cpp:108-111); ObjectToWorldNormal
    // is the row vector (n, 1) times model^-1 (SR/IShader.h:71-75)
    F3 n = normalize3(n_obj);
    UniformRow mi = u.model_inv;
    float wn2 = 0.0f; wn2 += mi[14] * 1.0f; wn2 += mi[10] * n.z; wn2 += mi[6] * n.y; wn2 += mi[2] * n.x;
    float wn1 = 0.0f; wn1 += mi[13] * 1.0f; wn1 += mi[9] * n.z; wn1 += mi[5] * n.y; wn1 += mi[1] * n.x;
    float wn0 = 0.0f; wn0 += mi[12] * 1.0f; wn0 += mi[8] * n.z; wn0 += mi[4] * n.y; wn0 += mi[0] * n.x;
    if (kind == SHADER_GROUND || kind == SHADER_TOON) {
        vary[0] = saturate(dot3(f3(wn0, wn1, wn2), f3(u.light_dir[0], u.light_dir[1], u.light_dir[2])));
        return;
    }
    if (kind == SHADER_TEXTURE_LIGHT) {
        vary[0] = wn0; vary[1] = wn1; vary[2] = wn2; vary[3] = tu; vary[4] = tv; vary[5] = vary[6] = vary[7] = 0.0f;
        return;
    }
    // BLINN / NORMALMAP: ObjectToWorldPos (SR/IShader.h:63-65)
    vary[2] = dot4(u.model + 8, p.x, p.y, p.z, 1.0f);
    vary[1] = dot4(u.model + 4, p.x, p.y, p.z, 1.0f);
    vary[0] = dot4(u.model + 0, p.x, p.y, p.z, 1.0f);
    vary[3] = wn0; vary[4] = wn1; vary[5] = wn2;
    vary[6] = tu; vary[7] = tv;
}

// IShader::is_in_shadow (SR/IShader.h:107-129).  The shadow map is an RGBA8
// render target whose R byte holds the quantised light-space depth; in the
// fp32 extension its fp32 depth buffer d gives the nearest caster's light clip
// z as 2*d - 1.
struct ShadowMap {
    const uint32_t* color;
    const float* depth;
    int w, h;
};
__device__ __forceinline__ int is_in_shadow(Uniforms u, const ShadowMap& sm, const float dp[4], float n_dot_l) {
    if (u.enable_shadow && sm.color) {
        float widht = (float)sm.w, height = (float)sm.h;
        float ndx = dp[0] / dp[3], ndy = dp[1] / dp[3];
        int iw = (int)widht, ih = (int)height;
        float posx = (ndx + 1) * 0.5f * (float)iw;
        float posy = (ndy + 1) * 0.5f * (float)ih;
        float depth_bias = 0.05f * (1 - n_dot_l);
        if (depth_bias < 0.005f) depth_bias = 0.01f;
        float current_depth = dp[2] - depth_bias;
        if (posx < 0 || posy < 0 || posx >= widht || posy >= height) return 1;
        int px = (int)posx, py = (int)posy;
        size_t texel = (size_t)py * (size_t)sm.w + (size_t)px;
        float closest_depth;
        if (u.shadow_fp32) closest_depth = 2.f * gld(sm.depth + texel) - 1.f;
        else closest_depth = byte_div255(gld(sm.color + texel) & 0xFF);
        return current_depth < closest_depth;
    }
    return 1;
}

// Shared tail of BlinnShader/NormalMapShader::fragment (SR/IShader.cpp:97-108, 150-161).
__device__ __forceinline__ C3 blinn_tail(Uniforms u, const Tex& diffuse, const ShadowMap& sm, F3 wp, F3 n,
                                         float tu, float tv) {
    F3 L = f3(u.light_dir[0], u.light_dir[1], u.light_dir[2]);
    C3 albedo = cmulc(tex_diffuse(diffuse, tu, tv), c3p(u.mat_color));
    C3 ambient = cmulc(c3p(u.ambient), albedo);
    float n_dot_l = saturate(dot3(n, L));
    C3 diff = cmulf(cmulc(c3p(u.light_color), albedo), n_dot_l);
    F3 view_dir = normalize3(sub3(f3(u.view_pos[0], u.view_pos[1], u.view_pos[2]), wp));
    F3 half_dir = normalize3(add3(view_dir, L));
    C3 spec = cmulf(cmulc(c3p(u.light_color), c3p(u.mat_spec)), powf(saturate(dot3(n, half_dir)), u.gloss));
    float dp[4];
    dp[3] = dot4(u.light_vp + 12, wp.x, wp.y, wp.z, 1.0f);
    dp[2] = dot4(u.light_vp + 8, wp.x, wp.y, wp.z, 1.0f);
    dp[1] = dot4(u.light_vp + 4, wp.x, wp.y, wp.z, 1.0f);
    dp[0] = dot4(u.light_vp + 0, wp.x, wp.y, wp.z, 1.0f);
    int shadow = is_in_shadow(u, sm, dp, n_dot_l);
    return cadd(ambient, cmulf(cadd(diff, spec), (float)shadow));
}

// ---- fragment programs (SR/IShader.cpp:11-180) on the interpolated varyings ----
__device__ __forceinline__ C3 fragment_program(int kind, Uniforms u, const float* vy, const Tex& diffuse,
                                               const Tex& normal, const ShadowMap& sm) {
    const C3 white = c3(1.f, 1.f, 1.f);
    switch (kind) {
    case SHADER_GROUND:
        return cmulf(white, vy[0]);
    case SHADER_TOON: {
        float intensity = vy[0];
        if (intensity > .85) intensity = 1;
        else if (intensity > .60) intensity = (float).80;
        else if (intensity > .45) intensity = (float).60;
        else if (intensity > .30) intensity = (float).45;
        else if (intensity > .15) intensity = (float).30;
        return cmulf(c3(1, 155 / 255.f, 0), intensity);
    }
    case SHADER_TEXTURE:
        return tex_diffuse(diffuse, vy[0], vy[1]);
    case SHADER_TEXTURE_LIGHT: {
        float intensity = saturate(dot3(f3(vy[0], vy[1], vy[2]), f3(u.light_dir[0], u.light_dir[1], u.light_dir[2])));
        return cmulf(tex_diffuse(diffuse, vy[3], vy[4]), intensity);
    }
    case SHADER_BLINN: {
        F3 n = normalize3(f3(vy[3], vy[4], vy[5]));
        return blinn_tail(u, diffuse, sm, f3(vy[0], vy[1], vy[2]), n, vy[6], vy[7]);
    }
    case SHADER_NORMALMAP: {
        float x = vy[3], y = vy[4], z = vy[5];
        F3 normal_v = f3(x, y, z);
        F3 t = f3(x * y / sqrtf(x * x + z * z), sqrtf(x * x + z * z), z * y / sqrtf(x * x + z * z));
        F3 b = cross3(normal_v, t);
        F3 bump = tex_normal(normal, vy[6], vy[7]);
        bump.x = bump.x * u.bump_scale;
        bump.y = bump.y * u.bump_scale;
        float bb = 0.0f;
        bb += bump.y * bump.y;
        bb += bump.x * bump.x;
        bump.z = (float)sqrt(1.0 - (double)saturate(bb));
        F3 tb;
        tb.z = dot3(f3(t.z, b.z, normal_v.z), bump);
        tb.y = dot3(f3(t.y, b.y, normal_v.y), bump);
        tb.x = dot3(f3(t.x, b.x, normal_v.x), bump);
        return blinn_tail(u, diffuse, sm, f3(vy[0], vy[1], vy[2]), normalize3(tb), vy[6], vy[7]);
    }
    case SHADER_SHADOW:
    default:
        return cmulf(white, vy[0]);
    }
}

// RenderBuffer::set_color: byte = (unsigned char)(c * 255) (SR/renderbuffer.cpp:38-44)
__device__ __forceinline__ uint32_t pack_rgb(C3 c, uint32_t alpha_word) {
    uint32_t r = (uint32_t)(c.r * 255), g = (uint32_t)(c.g * 255), b = (uint32_t)(c.b * 255);
    return (r & 0xFF) | ((g & 0xFF) << 8) | ((b & 0xFF) << 16) | (alpha_word & 0xFF000000u);
}

}  // namespace hana
