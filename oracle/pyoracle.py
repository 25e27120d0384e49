"""ctypes bindings for the CPU checkers (TEST INFRASTRUCTURE ONLY).

Two libraries expose the same entry points (hana_oracle.h):
  * liboracle.so          -- the C restatement, prefix ``ho_``;
  * _ref/libhanaref.so    -- the reference's own sources compiled by
                             oracle/Makefile, prefix ``ref_``;
  * _ref/libhanaref_dropin.so -- the same, with graphics_draw_triangle taken
                             from the MI355X drop-in (its ref_draw_model runs
                             the reference's Model/IShader/RenderBuffer through
                             libhana_gpu.so; every other entry point is CPU).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use them.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhanaref.so")
DROPIN_SO = os.path.join(HERE, "_ref", "libhanaref_dropin.so")

FLT_MAX = float(np.finfo(np.float32).max)


class Uniforms(C.Structure):
    _fields_ = [("model", C.c_float * 16), ("model_inv", C.c_float * 16), ("view", C.c_float * 16),
                ("proj", C.c_float * 16), ("light_vp", C.c_float * 16), ("camera_vp", C.c_float * 16),
                ("view_pos", C.c_float * 3), ("light_dir", C.c_float * 3), ("light_color", C.c_float * 4),
                ("ambient", C.c_float * 4), ("mat_color", C.c_float * 4), ("mat_specular", C.c_float * 4),
                ("gloss", C.c_float), ("bump_scale", C.c_float), ("enable_shadow", C.c_int32),
                ("shadow_fp32", C.c_int32)]


class A2V(C.Structure):
    _fields_ = [("obj_pos", C.c_float * 3), ("obj_normal", C.c_float * 3), ("uv", C.c_float * 2)]


class V2F(C.Structure):
    _fields_ = [("f", C.c_float * 13)]


class Image(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("bytespp", C.c_int32), ("pad_", C.c_int32),
                ("data", C.c_void_p)]


class Target(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("color", C.c_void_p), ("depth", C.c_void_p)]


class MeshS(C.Structure):
    _fields_ = [("nfaces", C.c_int32), ("flags", C.c_int32), ("pos", C.c_void_p), ("normals", C.c_void_p),
                ("uvs", C.c_void_p), ("idx_pos", C.c_void_p), ("idx_nrm", C.c_void_p), ("idx_uv", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("faces", C.c_int64), ("clipped_faces", C.c_int64), ("triangles", C.c_int64),
                ("culled", C.c_int64), ("bbox_pixels", C.c_int64), ("fragments", C.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class HostTarget:
    """RenderBuffer-shaped host arrays (RGBA8 + fp32 depth)."""

    def __init__(self, width: int, height: int, rgba=(0.0, 0.0, 0.0, 1.0), depth: float = FLT_MAX):
        self.width, self.height = int(width), int(height)
        self.color = np.zeros((self.height, self.width, 4), np.uint8)
        self.depth = np.zeros((self.height, self.width), np.float32)
        self.clear(rgba, depth)

    def clear(self, rgba=(0.0, 0.0, 0.0, 1.0), depth: float = FLT_MAX):
        b = []
        for v in rgba:
            f = np.float32(v) * np.float32(255)
            b.append(int(f) if 0 <= f < 256 else 255)
        self.color[...] = np.array(b, np.uint8)
        self.depth[...] = np.float32(depth)

    def struct(self) -> Target:
        return Target(self.width, self.height, _ptr(self.color), _ptr(self.depth))


def write_obj(mesh, path) -> None:
    """Wavefront OBJ the reference's parser reads (v / vt / vn, f v/t/n with
    one index for all three); %.9g round-trips every float32 exactly."""
    with open(path, "w") as f:
        for p in np.asarray(mesh.pos, np.float32):
            f.write("v %.9g %.9g %.9g\n" % tuple(float(x) for x in p))
        for t in np.asarray(mesh.uv, np.float32):
            f.write("vt %.9g %.9g 0\n" % tuple(float(x) for x in t))
        for n in np.asarray(mesh.nrm, np.float32):
            f.write("vn %.9g %.9g %.9g\n" % tuple(float(x) for x in n))
        for a, b, c in np.asarray(mesh.idx, np.int64) + 1:
            f.write(f"f {a}/{a}/{a} {b}/{b}/{b} {c}/{c}/{c}\n")


def image_struct(img: np.ndarray | None) -> Image:
    if img is None:
        return Image(0, 0, 0, 0, None)
    assert img.dtype == np.uint8 and img.flags.c_contiguous
    h, w = img.shape[:2]
    bpp = 1 if img.ndim == 2 else img.shape[2]
    return Image(w, h, bpp, 0, _ptr(img))


class CPURenderer:
    """The oracle (kind='port') or the compiled reference (kind='reference')."""

    def __init__(self, which: str = "oracle"):
        if which == "oracle":
            path, self.prefix, self.kind = ORACLE_SO, "ho_", "port"
        elif which == "reference":
            path, self.prefix, self.kind = REF_SO, "ref_", "reference"
        elif which == "dropin":
            path, self.prefix, self.kind = DROPIN_SO, "ref_", "dropin"
        else:
            raise ValueError(which)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run __graft_entry__.build())")
        self.path = path
        self.lib = C.CDLL(path)
        L, p = self.lib, self.prefix
        fp = C.POINTER(C.c_float)
        def sig(name, res, *args):
            f = getattr(L, p + name)
            f.restype = res
            f.argtypes = list(args)
            return f
        self._mat4_mul = sig("mat4_mul", None, fp, fp, fp)
        self._mat4_mul_vec4 = sig("mat4_mul_vec4", None, fp, fp, fp)
        self._invert = sig("mat4_invert", None, fp, fp)
        self._lookat = sig("lookat", None, fp, fp, fp, fp)
        self._ortho = sig("orthographic", None, C.c_float, C.c_float, C.c_float, C.c_float, fp)
        self._persp = sig("perspective", None, C.c_float, C.c_float, C.c_float, C.c_float, fp)
        self._rot = {a: sig("rotate_" + a, None, C.c_float, fp) for a in "xyz"}
        self._rotate = sig("rotate", None, C.c_float, C.c_float, C.c_float, C.c_float, fp)
        self._translate = sig("translate", None, C.c_float, C.c_float, C.c_float, fp)
        self._scale = sig("scale", None, C.c_float, C.c_float, C.c_float, fp)
        self._viewport = sig("viewport_transform", None, C.c_int, C.c_int, fp, fp)
        self._model_matrix = sig("model_matrix", None, fp, fp, fp, fp)
        self._build = sig("build_uniforms", None, fp, fp, C.c_float, fp, fp, fp, fp, C.c_int, C.c_int, C.c_int,
                          C.POINTER(Uniforms))
        self._bary = sig("barycentric", None, fp, fp, fp, fp, fp)
        self._clip = sig("clip_triangle", C.c_int, C.POINTER(V2F), C.POINTER(V2F))
        self._backface = sig("is_back_facing", C.c_int, fp)
        self._vertex = sig("vertex", V2F, C.c_int, C.POINTER(Uniforms), C.POINTER(A2V))
        self._fragment = sig("fragment", C.c_int, C.c_int, C.POINTER(Uniforms), C.POINTER(Image),
                             C.POINTER(Image), C.POINTER(Target), C.POINTER(V2F), fp)
        self._draw = sig("draw", C.c_int, C.POINTER(MeshS), C.c_int, C.POINTER(Uniforms), C.POINTER(Image),
                         C.POINTER(Image), C.POINTER(Target), C.POINTER(Target), C.POINTER(Stats),
                         C.c_void_p)
        self._frame = sig("render_frame", C.c_int, C.POINTER(MeshS), C.c_int, C.c_int, C.POINTER(Uniforms),
                          C.POINTER(Image), C.POINTER(Image), C.POINTER(Target), C.POINTER(Target),
                          C.POINTER(Stats), C.POINTER(Stats))
        self._camera_update = getattr(L, "ref_camera_update", None)
        if self._camera_update is not None:
            self._camera_update.restype = None
            self._camera_update.argtypes = [fp, fp, C.c_float, fp, fp, fp]
            self._scene_keys = L.ref_scene_keys
            self._scene_keys.restype = C.c_int
            self._scene_keys.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int, fp]
        self._draw_model = getattr(L, "ref_draw_model", None)
        if self._draw_model is not None:
            self._draw_model.restype = C.c_int
            self._draw_model.argtypes = [C.c_char_p, C.c_int, C.POINTER(Uniforms), C.POINTER(Image),
                                         C.POINTER(Image), C.POINTER(Target), C.POINTER(Target)]

        self._model_corners = getattr(L, "ref_model_corners", None)
        if self._model_corners is not None:
            self._model_corners.restype = C.c_int
            self._model_corners.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
            self._tga_read = L.ref_tga_read
            self._tga_read.restype = C.c_int
            self._tga_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.c_void_p, C.c_size_t]
            self._tga_write = L.ref_tga_write
            self._tga_write.restype = C.c_int
            self._tga_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int]

    # -- the reference's own OBJ / TGA code (reference library only) -----
    def model_corners(self, obj_path, max_faces=1 << 24):
        """Per-corner vert/normal/uv as graphics_draw_triangle fetches them
        from a Model read from obj_path: (F*3, 3), (F*3, 3), (F*3, 2)."""
        path = os.fsencode(obj_path)
        nf = self._model_corners(path, None, None, None, 0)
        pos = np.zeros((3 * nf, 3), np.float32)
        nrm = np.zeros((3 * nf, 3), np.float32)
        uv = np.zeros((3 * nf, 2), np.float32)
        self._model_corners(path, _ptr(pos), _ptr(nrm), _ptr(uv), min(nf, max_faces))
        return pos, nrm, uv

    def tga_read(self, path, flip_vertically=False, cap=1 << 26):
        w, h, b = C.c_int(), C.c_int(), C.c_int()
        buf = np.zeros(cap, np.uint8)
        ok = self._tga_read(os.fsencode(path), int(flip_vertically), C.byref(w), C.byref(h), C.byref(b), _ptr(buf), cap)
        if not ok:
            return None
        n = w.value * h.value * b.value
        shape = (h.value, w.value) if b.value == 1 else (h.value, w.value, b.value)
        return buf[:n].reshape(shape).copy()

    def tga_write(self, path, img, rle=True):
        a = np.ascontiguousarray(img, np.uint8)
        bpp = 1 if a.ndim == 2 else a.shape[2]
        return bool(self._tga_write(os.fsencode(path), a.shape[1], a.shape[0], bpp, _ptr(a), int(rle)))

    # -- small helpers ---------------------------------------------------
    @staticmethod
    def _f(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        return a, a.ctypes.data_as(C.POINTER(C.c_float))

    def _mat_out(self, fn, *args):
        out = np.zeros(16, np.float32)
        fn(*args, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out.reshape(4, 4)

    def mat4_mul(self, a, b):
        a, pa = self._f(a); b, pb = self._f(b)
        return self._mat_out(self._mat4_mul, pa, pb)

    def mat4_mul_vec4(self, m, v):
        m, pm = self._f(m); v, pv = self._f(v)
        out = np.zeros(4, np.float32)
        self._mat4_mul_vec4(pm, pv, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def invert(self, m):
        m, pm = self._f(m)
        return self._mat_out(self._invert, pm)

    def lookat(self, e, t, u):
        e, pe = self._f(e); t, pt = self._f(t); u, pu = self._f(u)
        return self._mat_out(self._lookat, pe, pt, pu)

    def orthographic(self, r, t, n, f):
        return self._mat_out(self._ortho, r, t, n, f)

    def perspective(self, fovy, aspect, n, f):
        return self._mat_out(self._persp, fovy, aspect, n, f)

    def rotate_axis(self, axis, angle):
        return self._mat_out(self._rot[axis], angle)

    def rotate(self, angle, vx, vy, vz):
        return self._mat_out(self._rotate, angle, vx, vy, vz)

    def translate(self, x, y, z):
        return self._mat_out(self._translate, x, y, z)

    def scale(self, x, y, z):
        return self._mat_out(self._scale, x, y, z)

    def viewport(self, w, h, ndc):
        ndc, p = self._f(ndc)
        out = np.zeros(3, np.float32)
        self._viewport(w, h, p, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def model_matrix(self, pos, rot, scl):
        a, pa = self._f(pos); b, pb = self._f(rot); c, pc = self._f(scl)
        return self._mat_out(self._model_matrix, pa, pb, pc)

    def barycentric(self, A, B, C_, P):
        arrs = [self._f(x) for x in (A, B, C_, P)]
        out = np.zeros(3, np.float32)
        self._bary(*[p for _, p in arrs], out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def clip_triangle(self, v2f3: np.ndarray):
        """v2f3: (3,13) float32 -> (n,13) float32."""
        inp = (V2F * 3)()
        C.memmove(inp, np.ascontiguousarray(v2f3, np.float32).ctypes.data, 3 * 13 * 4)
        out = (V2F * 10)()
        n = self._clip(inp, out)
        return np.frombuffer(bytes(out), np.float32).reshape(10, 13)[:n].copy()

    def is_back_facing(self, ndc9):
        a, p = self._f(ndc9)
        return bool(self._backface(p))

    def camera_update(self, pos, target, aspect, motion):
        """Camera::update_transform (reference library only): new (position, target)."""
        if self._camera_update is None:
            raise RuntimeError(f"{self.path} has no camera entry point")
        fp = C.POINTER(C.c_float)
        a = [np.ascontiguousarray(x, np.float32) for x in (pos, target, motion)]
        op, ot = np.zeros(3, np.float32), np.zeros(3, np.float32)
        self._camera_update(*(x.ctypes.data_as(fp) for x in a[:2]), np.float32(aspect), a[2].ctypes.data_as(fp),
                            op.ctypes.data_as(fp), ot.ctypes.data_as(fp))
        return op, ot

    def scene_keys(self, obj_path, width, height, keys):
        """SingleModelScene driven by pressed keys (reference library only):
        per key (light x, y, z, enable_shadow, shader index)."""
        if self._camera_update is None:
            raise RuntimeError(f"{self.path} has no scene entry point")
        k = (C.c_int * len(keys))(*keys)
        out = np.zeros((len(keys), 5), np.float32)
        rc = self._scene_keys(str(obj_path).encode(), width, height, k, len(keys),
                              out.ctypes.data_as(C.POINTER(C.c_float)))
        if rc:
            raise RuntimeError("ref_scene_keys: shader line not recognised")
        return out

    def build_uniforms(self, scene, enable_shadow=None) -> Uniforms:
        u = Uniforms()
        vals = [self._f(np.array(x, np.float32)) for x in (scene.cam_pos, scene.cam_target)]
        light = self._f(np.array(scene.light_pos, np.float32))
        op = self._f(np.array(scene.obj_position, np.float32))
        orot = self._f(np.array(scene.obj_rotation, np.float32))
        osc = self._f(np.array(scene.obj_scale, np.float32))
        es = scene.shadow if enable_shadow is None else enable_shadow
        self._build(vals[0][1], vals[1][1], np.float32(scene.width) / np.float32(scene.height), light[1],
                    op[1], orot[1], osc[1], scene.width, scene.height, int(es), C.byref(u))
        u.shadow_fp32 = int(getattr(scene, "shadow_fp32", False))
        return u

    def vertex(self, kind, u: Uniforms, a2v: np.ndarray) -> np.ndarray:
        a = A2V()
        C.memmove(C.byref(a), np.ascontiguousarray(a2v, np.float32).ctypes.data, 32)
        r = self._vertex(kind, C.byref(u), C.byref(a))
        return np.frombuffer(bytes(r), np.float32).copy()

    def fragment(self, kind, u: Uniforms, v2f: np.ndarray, diffuse=None, normal=None, shadow: HostTarget | None = None):
        v = V2F()
        C.memmove(C.byref(v), np.ascontiguousarray(v2f, np.float32).ctypes.data, 52)
        d, n = image_struct(diffuse), image_struct(normal)
        s = shadow.struct() if shadow is not None else None
        out = np.zeros(4, np.float32)
        self._fragment(kind, C.byref(u), C.byref(d), C.byref(n), C.byref(s) if s is not None else None,
                       C.byref(v), out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def _mesh(self, mesh):
        keep = [np.ascontiguousarray(mesh.pos, np.float32), np.array(mesh.nrm, np.float32, copy=True),
                np.ascontiguousarray(mesh.uv, np.float32), np.ascontiguousarray(mesh.idx, np.int32)]
        flags = 1 if getattr(mesh, "fetched_normals", False) else 0     # HO_MESH_FETCHED_NORMALS
        m = MeshS(mesh.nfaces, flags, _ptr(keep[0]), _ptr(keep[1]), _ptr(keep[2]), _ptr(keep[3]), None, None)
        return m, keep

    def draw(self, mesh, kind, u: Uniforms, target: HostTarget, diffuse=None, normal=None,
             shadow: HostTarget | None = None, winner: bool = False):
        m, keep = self._mesh(mesh)
        d, n = image_struct(diffuse), image_struct(normal)
        st = Stats()
        win = np.full((target.height, target.width), -1, np.int32) if winner else None
        t = target.struct()
        s = shadow.struct() if shadow is not None else None
        self._draw(C.byref(m), kind, C.byref(u), C.byref(d), C.byref(n), C.byref(s) if s is not None else None,
                   C.byref(t), C.byref(st), _ptr(win) if win is not None else None)
        return st.as_dict(), win

    def draw_model(self, obj_path: str, kind, u: Uniforms, target: HostTarget, diffuse=None, normal=None,
                   shadow: HostTarget | None = None) -> int:
        """graphics_draw_triangle itself on a Model read from obj_path by the
        reference's OBJ parser (reference and drop-in libraries only)."""
        if self._draw_model is None:
            raise RuntimeError(f"{self.path} has no draw_model entry point")
        d, n = image_struct(diffuse), image_struct(normal)
        t = target.struct()
        s = shadow.struct() if shadow is not None else None
        return self._draw_model(os.fsencode(obj_path), kind, C.byref(u), C.byref(d), C.byref(n),
                                C.byref(s) if s is not None else None, C.byref(t))

    def render_frame(self, scene, u: Uniforms | None = None, target: HostTarget | None = None,
                     shadow_map: HostTarget | None = None):
        """DrawModel::draw on a cleared target (Black, FLT_MAX); returns
        (target, shadow_map_before_clear_is_not_kept, stats_shadow, stats_main)."""
        if u is None:
            u = self.build_uniforms(scene)
        if target is None:
            target = HostTarget(scene.width, scene.height)
        if shadow_map is None and scene.shadow:
            sw, sh = scene.shadow_wh
            shadow_map = HostTarget(sw, sh)
        m, keep = self._mesh(scene.mesh)
        d, n = image_struct(scene.diffuse), image_struct(scene.normal)
        st_s, st_m = Stats(), Stats()
        t = target.struct()
        s = shadow_map.struct() if shadow_map is not None else None
        self._frame(C.byref(m), scene.shader, int(scene.shadow), C.byref(u), C.byref(d), C.byref(n),
                    C.byref(s) if s is not None else None, C.byref(t), C.byref(st_s), C.byref(st_m))
        return target, st_s.as_dict(), st_m.as_dict()

    def render_passes(self, scene, u: Uniforms | None = None):
        """Shadow pass then main pass, keeping the shadow map (no final clear)."""
        if u is None:
            u = self.build_uniforms(scene)
        target = HostTarget(scene.width, scene.height)
        shadow_map = None
        st_s = None
        if scene.shadow:
            sw, sh = scene.shadow_wh
            shadow_map = HostTarget(sw, sh)
            st_s, _ = self.draw(scene.mesh, 6, u, shadow_map, scene.diffuse, scene.normal)
        st_m, _ = self.draw(scene.mesh, scene.shader, u, target, scene.diffuse, scene.normal, shadow_map)
        return target, shadow_map, st_s, st_m
