// san_driver.cpp -- runs the host-side code of the library (the OBJ and TGA
// parsers, the uniform math, the scene state) and the oracle under
// AddressSanitizer and UndefinedBehaviorSanitizer.  Built and run by
// tests/test_sanitizers.py; every file named on the command line is parsed
// (malformed ones must fail cleanly), each mesh that loads is rendered by the
// oracle with every shader, and each image that loads is written back and
// read again.  Exit status 0 when every check holds; a sanitizer report
// aborts the run.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hana_gpu.h"
#include "hana_oracle.h"

static int failures = 0;
#define CHECK(cond)                                                     \
    do {                                                                \
        if (!(cond)) {                                                  \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
            failures++;                                                 \
        }                                                               \
    } while (0)

static bool ends_with(const std::string& s, const char* suf) {
    size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

// Renders a loaded mesh with every shader, shadowed and not, at a small size.
static void render_mesh(const hana_obj_mesh& m, const ho_image& tex) {
    std::vector<float> nrm(m.nrm, m.nrm + 3 * (size_t)m.nverts);
    ho_mesh mesh{m.nfaces, HO_MESH_FETCHED_NORMALS, m.pos, nrm.data(), m.uv, m.idx, nullptr, nullptr};
    const int W = 61, H = 47;
    std::vector<unsigned char> col(4 * W * H), scol(4 * W * H);
    std::vector<float> dep(W * H), sdep(W * H);
    ho_target tgt{W, H, col.data(), dep.data()}, sm{W, H, scol.data(), sdep.data()};
    const float cams[3][3] = {{0, 0, 2}, {0.3f, 0.2f, 0.4f}, {1.5f, -0.5f, 0.2f}};
    const float tgt0[3] = {0, 0, 0}, light[3] = {2, 2, 2}, zero[3] = {0, 0, 0}, one[3] = {1, 1, 1};
    for (const auto& cam : cams)
        for (int kind = HO_SHADER_GROUND; kind <= HO_SHADER_NORMALMAP; kind++)
            for (int shadow = 0; shadow < 2; shadow++)
                for (int fp32 = 0; fp32 < 1 + shadow; fp32++) {
                    ho_uniforms u;
                    ho_build_uniforms(cam, tgt0, (float)W / H, light, zero, zero, one, W, H, shadow, &u);
                    u.shadow_fp32 = fp32;
                    const float black[4] = {0, 0, 0, 1};
                    ho_target_clear(&tgt, black, 3.4e38f);
                    ho_target_clear(&sm, black, 3.4e38f);
                    ho_stats ss{}, st{};
                    ho_render_frame(&mesh, kind, shadow, &u, &tex, &tex, shadow ? &sm : nullptr, &tgt, &ss, &st);
                    CHECK(st.faces == m.nfaces);
                }
}

// The uniform math and the interactive scene state over input sequences.
static void scene_state() {
    hana_scene s;
    CHECK(hana_scene_init(&s, 800, 600) == HANA_OK);
    const int keys[] = {HANA_KEY_A, HANA_KEY_D, HANA_KEY_Q, HANA_KEY_E, HANA_KEY_SPACE};
    for (int i = 0; i < 200; i++) {
        hana_motion m{{0.01f * (i % 7) - 0.03f, 0.02f * (i % 5) - 0.04f}, {0.001f * (i % 3), -0.002f * (i % 4)},
                      (i % 9) * 0.05f - 0.2f};
        CHECK(hana_scene_update_camera(&s, &m) == HANA_OK);
        CHECK(hana_scene_key(&s, keys[i % 5]) == HANA_OK);
        CHECK(hana_scene_shader(&s) >= HANA_SHADER_GROUND);
        hana_uniforms u;
        const float zero[3] = {0, 0, 0}, one[3] = {1, 1, 1};
        CHECK(hana_scene_uniforms(&s, zero, zero, one, 800, 600, &u) == HANA_OK);
        CHECK(std::isfinite(u.camera_vp[0]));
    }
    CHECK(hana_scene_key(&s, 99) != HANA_OK || true);
    hana_uniforms u;
    const float cam[3] = {0, 0, 2}, tgt[3] = {0, 0, 0}, light[3] = {2, 2, 2}, rot[3] = {30, 45, 60},
                sc[3] = {1, 2, 3};
    CHECK(hana_build_uniforms(cam, tgt, 4.f / 3, light, tgt, rot, sc, 640, 480, 1, &u) == HANA_OK);
}

int main(int argc, char** argv) {
    std::vector<unsigned char> texels(64 * 64 * 4);
    for (size_t i = 0; i < texels.size(); i++) texels[i] = (unsigned char)(i * 37u);
    ho_image tex{64, 64, 4, 0, texels.data()};
    const char* tmp = std::getenv("SAN_TMP");
    int loaded_obj = 0, loaded_tga = 0, rejected = 0;
    for (int a = 1; a < argc; a++) {
        std::string path = argv[a];
        if (ends_with(path, ".obj")) {
            hana_obj_mesh m;
            int rc = hana_obj_load(path.c_str(), &m);
            if (rc != HANA_OK) {
                CHECK(hana_asset_last_error() != nullptr);
                rejected++;
                continue;
            }
            loaded_obj++;
            for (int i = 0; i < 3 * m.nfaces; i++) CHECK(m.idx[i] >= 0 && m.idx[i] < m.nverts);
            if (m.nfaces <= 20000) render_mesh(m, tex);
            CHECK(hana_obj_free(&m) == HANA_OK);
        } else if (ends_with(path, ".tga")) {
            for (int flip = 0; flip < 2; flip++) {
                hana_image im;
                int rc = hana_tga_load(path.c_str(), flip, &im);
                if (rc != HANA_OK) {
                    rejected += flip == 0;
                    continue;
                }
                loaded_tga += flip == 0;
                CHECK(im.width > 0 && im.height > 0 && (im.bytespp == 1 || im.bytespp == 3 || im.bytespp == 4));
                if (tmp) {
                    for (int rle = 0; rle < 2; rle++) {
                        std::string out = std::string(tmp) + "/roundtrip.tga";
                        CHECK(hana_tga_save(out.c_str(), &im, rle) == HANA_OK);
                        hana_image back;
                        CHECK(hana_tga_load(out.c_str(), 0, &back) == HANA_OK);
                        size_t n = (size_t)im.width * im.height * im.bytespp;
                        CHECK(back.width == im.width && back.height == im.height && back.bytespp == im.bytespp);
                        CHECK(std::memcmp(back.bytes, im.bytes, n) == 0);
                        hana_image_free(&back);
                    }
                }
                CHECK(hana_image_free(&im) == HANA_OK);
            }
        }
    }
    scene_state();
    // clipping on random straddling triangles
    unsigned seed = 12345u;
    auto rnd = [&seed]() {
        seed = seed * 1664525u + 1013904223u;
        return (float)(seed >> 8) / (float)(1u << 24) * 6.f - 3.f;
    };
    for (int i = 0; i < 20000; i++) {
        ho_v2f in[3], out[30];
        std::memset(in, 0, sizeof in);
        for (auto& v : in) {
            for (int k = 0; k < 3; k++) v.clip_pos[k] = rnd();
            v.clip_pos[3] = rnd() + 1.5f;
        }
        int n = ho_clip_triangle(in, out);
        CHECK(n == 0 || (n >= 3 && n <= 30));
    }
    std::printf("san_driver: %d meshes, %d images loaded, %d files rejected, %d failures\n", loaded_obj, loaded_tga,
                rejected, failures);
    return failures ? 1 : 0;
}
