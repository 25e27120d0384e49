// hana_hostmath.h -- host-side uniform math with the reference's fp32
// operation order (SR/matrix.h, SR/maths.cpp, SR/gameobject.cpp, SR/scene.h).
// Built with -ffp-contract=off so that every product and sum rounds as the
// reference's do; the uniforms it produces are bit-identical to the
// reference's ShaderData for the same camera, light and transform.
#pragma once
#include "hana_types.h"

struct hana_uniforms;

namespace hana {

void m4_mul(const float* a, const float* b, float* out);        // Matrix*Matrix (SR/matrix.h:118-123)
void m4_invert(const float* m, float* out);                     // Matrix::invert (SR/matrix.h:90-109)
void m4_lookat(const float* eye, const float* target, const float* up, float* out);   // SR/maths.cpp:180-195
void m4_perspective(float fovy, float aspect, float n, float f, float* out);         // SR/maths.cpp:243-255
void m4_orthographic(float right, float top, float n, float f, float* out);          // SR/maths.cpp:214-223
void m4_model(const float* pos, const float* rot_deg, const float* scale, float* out);  // SR/gameobject.cpp:12-17

void build_uniforms(const float cam_pos[3], const float cam_target[3], float cam_aspect, const float light_pos[3],
                    const float obj_position[3], const float obj_rotation_deg[3], const float obj_scale[3], int fb_w,
                    int fb_h, int enable_shadow, hana_uniforms* out);

// Per-pass device uniforms: folds camera_vp*model and light_vp*model once.
void derive_dev_uniforms(const hana_uniforms& u, DevUniforms& d);

}  // namespace hana
