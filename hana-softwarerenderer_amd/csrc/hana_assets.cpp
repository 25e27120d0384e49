// hana_assets.cpp -- the renderer's on-disk formats on the host: Wavefront OBJ
// as Model reads it (SR/model.cpp:6-48) and TGA as TGAImage reads and writes
// it (SR/tgaimage.cpp:40-246), producing the arrays hana_mesh_create and
// hana_texture_create take.  No device code.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hana_gpu.h"

namespace {

thread_local std::string g_error;

int asset_fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

struct V3 {
    float v[3];
};

// Vector3f::normalize: s = 1/sqrtf((x*x + y*y) + z*z), v * s (SR/vector.h:41-42).
V3 renormalised(const V3& a) {
    float n = std::sqrt((a.v[0] * a.v[0] + a.v[1] * a.v[1]) + a.v[2] * a.v[2]);
    float s = 1.0f / n;
    return V3{{a.v[0] * s, a.v[1] * s, a.v[2] * s}};
}

// istream >> float in the "C" locale: a strtof of the next whitespace-delimited
// token prefix; a failed read leaves 0 and fails every later read of the line.
struct LineReader {
    const char* p;
    bool ok = true;
    explicit LineReader(const char* s) : p(s) {}
    void skip_ws() {
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == '\v' || *p == '\f') p++;
    }
    char get_char() {           // >> char: the next non-blank character
        if (!ok) return 0;
        skip_ws();
        if (!*p) {
            ok = false;
            return 0;
        }
        return *p++;
    }
    float get_float() {
        if (!ok) return 0.0f;
        skip_ws();
        char* end = nullptr;
        errno = 0;
        float f = std::strtof(p, &end);
        if (end == p) {
            ok = false;
            return 0.0f;
        }
        p = end;
        return f;
    }
    int get_int() {
        if (!ok) return 0;
        skip_ws();
        char* end = nullptr;
        long v = std::strtol(p, &end, 10);
        if (end == p) {
            ok = false;
            return 0;
        }
        p = end;
        return (int)v;
    }
};

struct Corner {
    int v, t, n;
};

#pragma pack(push, 1)
struct TgaHeader {   // the 18-byte TGA file header, little-endian
    uint8_t idlength, colormaptype, datatypecode;
    int16_t colormaporigin, colormaplength;
    uint8_t colormapdepth;
    int16_t x_origin, y_origin, width, height;
    uint8_t bitsperpixel, imagedescriptor;
};
#pragma pack(pop)
static_assert(sizeof(TgaHeader) == 18, "TGA header is 18 bytes");

void flip_rows(uint8_t* data, int w, int h, int bpp) {
    size_t line = (size_t)w * bpp;
    std::vector<uint8_t> tmp(line);
    for (int j = 0; j < h / 2; j++) {
        uint8_t* a = data + (size_t)j * line;
        uint8_t* b = data + (size_t)(h - 1 - j) * line;
        std::memcpy(tmp.data(), a, line);
        std::memcpy(a, b, line);
        std::memcpy(b, tmp.data(), line);
    }
}

void flip_columns(uint8_t* data, int w, int h, int bpp) {
    for (int j = 0; j < h; j++) {
        uint8_t* row = data + (size_t)j * w * bpp;
        for (int i = 0; i < w / 2; i++) {
            uint8_t t[4];
            std::memcpy(t, row + (size_t)i * bpp, bpp);
            std::memcpy(row + (size_t)i * bpp, row + (size_t)(w - 1 - i) * bpp, bpp);
            std::memcpy(row + (size_t)(w - 1 - i) * bpp, t, bpp);
        }
    }
}

}  // namespace

extern "C" {

const char* hana_asset_last_error(void) { return g_error.c_str(); }

int hana_obj_load(const char* path, hana_obj_mesh* out) {
    if (!path || !out) return asset_fail(HANA_E_INVALID, "hana_obj_load: null argument");
    std::memset(out, 0, sizeof *out);
    std::ifstream in(path);
    if (!in) return asset_fail(HANA_E_INVALID, std::string("hana_obj_load: cannot open ") + path);
    std::vector<V3> verts, norms;
    std::vector<float> uvs;
    std::vector<Corner> corners;        // the first three corners of each face
    std::string line;
    int64_t faces = 0;
    while (std::getline(in, line)) {
        LineReader r(line.c_str());
        if (!line.compare(0, 2, "v ")) {
            r.get_char();
            V3 v{{0, 0, 0}};
            for (int i = 0; i < 3; i++) v.v[i] = r.get_float();
            verts.push_back(v);
        } else if (!line.compare(0, 3, "vn ")) {
            r.get_char();
            r.get_char();
            V3 n{{0, 0, 0}};
            for (int i = 0; i < 3; i++) n.v[i] = r.get_float();
            norms.push_back(n);
        } else if (!line.compare(0, 3, "vt ")) {
            r.get_char();
            r.get_char();
            float u = r.get_float();
            float v = r.get_float();
            uvs.push_back(u);
            uvs.push_back(v);
        } else if (!line.compare(0, 2, "f ")) {
            r.get_char();
            int n = 0;
            while (true) {   // v/vt/vn triplets, 1-based
                Corner c;
                c.v = r.get_int();
                r.get_char();
                c.t = r.get_int();
                r.get_char();
                c.n = r.get_int();
                if (!r.ok) break;
                if (n < 3) corners.push_back(Corner{c.v - 1, c.t - 1, c.n - 1});
                n++;
            }
            if (n < 3)
                return asset_fail(HANA_E_INVALID, "hana_obj_load: face " + std::to_string(faces) +
                                                      " has fewer than three v/vt/vn corners");
            faces++;
        }
    }
    if (faces > (1 << 27)) return asset_fail(HANA_E_INVALID, "hana_obj_load: at most 2^27 faces");
    int64_t nv = (int64_t)verts.size(), nt = (int64_t)uvs.size() / 2, nn = (int64_t)norms.size();
    for (size_t i = 0; i < corners.size(); i++) {
        const Corner& c = corners[i];
        if (c.v < 0 || c.v >= nv || c.t < 0 || c.t >= nt || c.n < 0 || c.n >= nn)
            return asset_fail(HANA_E_INVALID, "hana_obj_load: face " + std::to_string(i / 3) +
                                                  " indexes past the vertex, uv or normal list");
    }
    // The face loop fetches corners in order and Model::normal renormalises the
    // stored normal in place at each fetch (SR/model.cpp:108-111): the value a
    // corner sees is the stored vector after every earlier fetch of it.  One
    // device vertex per distinct (position, uv, fetched normal).
    struct Key {
        int v, t;
        uint32_t n[3];
        bool operator==(const Key& o) const {
            return v == o.v && t == o.t && n[0] == o.n[0] && n[1] == o.n[1] && n[2] == o.n[2];
        }
    };
    struct KeyHash {
        size_t operator()(const Key& k) const {
            uint64_t h = (uint64_t)(uint32_t)k.v * 0x9E3779B97F4A7C15ull ^ (uint64_t)(uint32_t)k.t * 0xC2B2AE3D27D4EB4Full;
            for (int i = 0; i < 3; i++) h = (h ^ k.n[i]) * 0x100000001B3ull;
            return (size_t)(h ^ (h >> 29));
        }
    };
    std::unordered_map<Key, int32_t, KeyHash> seen;
    seen.reserve(corners.size());
    std::vector<float> pos, nrm, uv;
    std::vector<int32_t> idx(corners.size());
    for (size_t i = 0; i < corners.size(); i++) {
        const Corner& c = corners[i];
        norms[c.n] = renormalised(norms[c.n]);
        const V3& n = norms[c.n];
        Key k{c.v, c.t, {0, 0, 0}};
        std::memcpy(k.n, n.v, sizeof k.n);
        auto it = seen.find(k);
        if (it == seen.end()) {
            int32_t id = (int32_t)(pos.size() / 3);
            it = seen.emplace(k, id).first;
            pos.insert(pos.end(), verts[c.v].v, verts[c.v].v + 3);
            nrm.insert(nrm.end(), n.v, n.v + 3);
            uv.push_back(uvs[2 * (size_t)c.t]);
            uv.push_back(uvs[2 * (size_t)c.t + 1]);
        }
        idx[i] = it->second;
    }
    out->nverts = (int32_t)(pos.size() / 3);
    out->nfaces = (int32_t)faces;
    size_t V = pos.size() / 3;
    out->pos = (float*)std::malloc(sizeof(float) * 3 * (V ? V : 1));
    out->nrm = (float*)std::malloc(sizeof(float) * 3 * (V ? V : 1));
    out->uv = (float*)std::malloc(sizeof(float) * 2 * (V ? V : 1));
    out->idx = (int32_t*)std::malloc(sizeof(int32_t) * (idx.empty() ? 1 : idx.size()));
    if (!out->pos || !out->nrm || !out->uv || !out->idx) {
        hana_obj_free(out);
        return asset_fail(HANA_E_NOMEM, "hana_obj_load: out of host memory");
    }
    // an empty list's data() may be null, which memcpy does not accept
    if (!pos.empty()) std::memcpy(out->pos, pos.data(), sizeof(float) * pos.size());
    if (!nrm.empty()) std::memcpy(out->nrm, nrm.data(), sizeof(float) * nrm.size());
    if (!uv.empty()) std::memcpy(out->uv, uv.data(), sizeof(float) * uv.size());
    if (!idx.empty()) std::memcpy(out->idx, idx.data(), sizeof(int32_t) * idx.size());
    return HANA_OK;
}

int hana_obj_free(hana_obj_mesh* m) {
    if (!m) return HANA_E_INVALID;
    std::free(m->pos);
    std::free(m->nrm);
    std::free(m->uv);
    std::free(m->idx);
    std::memset(m, 0, sizeof *m);
    return HANA_OK;
}

int hana_tga_load(const char* path, int flip_vertically, hana_image* out) {
    if (!path || !out) return asset_fail(HANA_E_INVALID, "hana_tga_load: null argument");
    std::memset(out, 0, sizeof *out);
    std::ifstream in(path, std::ios::binary);
    if (!in) return asset_fail(HANA_E_INVALID, std::string("hana_tga_load: cannot open ") + path);
    TgaHeader h;
    in.read(reinterpret_cast<char*>(&h), sizeof h);
    if (!in) return asset_fail(HANA_E_INVALID, "hana_tga_load: short header");
    int w = h.width, ht = h.height, bpp = h.bitsperpixel >> 3;
    if (w <= 0 || ht <= 0 || (bpp != 1 && bpp != 3 && bpp != 4))
        return asset_fail(HANA_E_INVALID, "hana_tga_load: bad bpp (or width/height) value");
    // the pixel data follows the header directly, as TGAImage reads it (the
    // image-id field and colour map are not skipped)
    size_t npix = (size_t)w * ht, nbytes = npix * bpp;
    std::vector<uint8_t> data(nbytes);
    if (h.datatypecode == 2 || h.datatypecode == 3) {
        in.read(reinterpret_cast<char*>(data.data()), (std::streamsize)nbytes);
        if (!in) return asset_fail(HANA_E_INVALID, "hana_tga_load: short pixel data");
    } else if (h.datatypecode == 10 || h.datatypecode == 11) {
        // run-length packets: header < 128 -> header+1 literal pixels,
        // else header-127 copies of one pixel
        size_t pix = 0;
        while (pix < npix) {
            int c = in.get();
            if (c == EOF) return asset_fail(HANA_E_INVALID, "hana_tga_load: truncated RLE data");
            bool literal = c < 128;
            size_t count = literal ? (size_t)c + 1 : (size_t)c - 127;
            if (pix + count > npix) return asset_fail(HANA_E_INVALID, "hana_tga_load: RLE data runs past the image");
            if (literal) {
                in.read(reinterpret_cast<char*>(data.data() + pix * bpp), (std::streamsize)(count * bpp));
                if (!in) return asset_fail(HANA_E_INVALID, "hana_tga_load: truncated RLE data");
            } else {
                uint8_t px[4];
                in.read(reinterpret_cast<char*>(px), bpp);
                if (!in) return asset_fail(HANA_E_INVALID, "hana_tga_load: truncated RLE data");
                for (size_t k = 0; k < count; k++) std::memcpy(data.data() + (pix + k) * bpp, px, bpp);
            }
            pix += count;
        }
    } else {
        return asset_fail(HANA_E_INVALID, "hana_tga_load: unknown data type " + std::to_string(h.datatypecode));
    }
    // rows end up top-down unless the descriptor says the file already is;
    // a right-to-left file is mirrored (SR/tgaimage.cpp:85-90)
    bool flip = !(h.imagedescriptor & 0x20);
    if (flip_vertically) flip = !flip;     // Model::load_texture flips once more (SR/model.cpp:80-81)
    if (flip) flip_rows(data.data(), w, ht, bpp);
    if (h.imagedescriptor & 0x10) flip_columns(data.data(), w, ht, bpp);
    out->bytes = (uint8_t*)std::malloc(nbytes);
    if (!out->bytes) return asset_fail(HANA_E_NOMEM, "hana_tga_load: out of host memory");
    std::memcpy(out->bytes, data.data(), nbytes);
    out->width = w;
    out->height = ht;
    out->bytespp = bpp;
    return HANA_OK;
}

int hana_tga_save(const char* path, const hana_image* img, int rle) {
    if (!path || !img || !img->bytes || img->width <= 0 || img->height <= 0 || img->width > 32767 ||
        img->height > 32767 || (img->bytespp != 1 && img->bytespp != 3 && img->bytespp != 4))
        return asset_fail(HANA_E_INVALID, "hana_tga_save: bad image");
    const int bpp = img->bytespp;
    const size_t npix = (size_t)img->width * img->height;
    const uint8_t* d = img->bytes;
    std::vector<uint8_t> buf;
    buf.reserve(18 + npix * bpp + 26);
    TgaHeader h;
    std::memset(&h, 0, sizeof h);
    h.bitsperpixel = (uint8_t)(bpp << 3);
    h.width = (int16_t)img->width;
    h.height = (int16_t)img->height;
    h.datatypecode = bpp == 1 ? (rle ? 11 : 3) : (rle ? 10 : 2);
    h.imagedescriptor = 0x20;   // rows stored top-down
    const uint8_t* hp = reinterpret_cast<const uint8_t*>(&h);
    buf.insert(buf.end(), hp, hp + sizeof h);
    if (!rle) {
        buf.insert(buf.end(), d, d + npix * bpp);
    } else {
        // Packets of at most 128 pixels.  A packet starting with two equal
        // pixels is a run over all the following equal pixels; otherwise it is
        // literal and ends before the first pair of equal neighbours, which
        // starts the next packet (SR/tgaimage.cpp:203-246).
        auto same = [&](size_t a, size_t b) { return std::memcmp(d + a * bpp, d + b * bpp, bpp) == 0; };
        size_t p = 0;
        while (p < npix) {
            size_t n = 1;
            bool run = p + 1 < npix && same(p, p + 1);
            if (run) {
                while (n < 128 && p + n < npix && same(p + n - 1, p + n)) n++;
                buf.push_back((uint8_t)(n + 127));
                buf.insert(buf.end(), d + p * bpp, d + (p + 1) * bpp);
            } else {
                while (n < 128 && p + n < npix) {
                    if (same(p + n - 1, p + n)) {
                        n--;
                        break;
                    }
                    n++;
                }
                buf.push_back((uint8_t)(n - 1));
                buf.insert(buf.end(), d + p * bpp, d + (p + n) * bpp);
            }
            p += n;
        }
    }
    // developer and extension area offsets (none), then the TGA 2.0 signature
    static const uint8_t tail[26] = {0, 0, 0, 0, 0, 0, 0, 0, 'T', 'R', 'U', 'E', 'V', 'I', 'S', 'I', 'O',
                                     'N', '-', 'X', 'F', 'I', 'L', 'E', '.', 0};
    buf.insert(buf.end(), tail, tail + sizeof tail);
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return asset_fail(HANA_E_INVALID, std::string("hana_tga_save: cannot create ") + path);
    size_t wrote = std::fwrite(buf.data(), 1, buf.size(), f);
    int closed = std::fclose(f);
    if (wrote != buf.size() || closed != 0) return asset_fail(HANA_E_INVALID, "hana_tga_save: write failed");
    return HANA_OK;
}

int hana_image_free(hana_image* img) {
    if (!img) return HANA_E_INVALID;
    std::free(img->bytes);
    std::memset(img, 0, sizeof *img);
    return HANA_OK;
}

}  // extern "C"
