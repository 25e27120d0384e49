"""ctypes binding of libhana_gpu.so (include/hana_gpu.h).

The product path is the HIP library only: if libhana_gpu.so is missing or no
GPU is visible, loading fails with an exception -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libhana_gpu.so")

HANA_OK, HANA_E_INVALID, HANA_E_DEVICE, HANA_E_NOMEM, HANA_E_OVERFLOW, HANA_E_UNSUPPORTED = 0, -1, -2, -3, -4, -5
GROUND, TOON, TEXTURE, TEXTURE_LIGHT, BLINN, NORMALMAP, SHADOW = range(7)
FLT_MAX = float(np.finfo(np.float32).max)
NO_WINNER = 0xFFFFFFFF

# C declarations, in header order: (name, restype, argtypes)
_P = C.c_void_p
_FP = C.POINTER(C.c_float)


class Uniforms(C.Structure):
    _fields_ = [("model", C.c_float * 16), ("model_inv", C.c_float * 16), ("view", C.c_float * 16),
                ("proj", C.c_float * 16), ("light_vp", C.c_float * 16), ("camera_vp", C.c_float * 16),
                ("view_pos", C.c_float * 3), ("light_dir", C.c_float * 3), ("light_color", C.c_float * 4),
                ("ambient", C.c_float * 4), ("mat_color", C.c_float * 4), ("mat_specular", C.c_float * 4),
                ("gloss", C.c_float), ("bump_scale", C.c_float), ("enable_shadow", C.c_int32),
                ("shadow_fp32", C.c_int32)]


class DrawDesc(C.Structure):
    _fields_ = [("mesh", _P), ("shader", C.c_int32), ("pad_", C.c_int32), ("uniforms", C.POINTER(Uniforms)),
                ("diffuse", _P), ("normal", _P), ("shadow_map", _P), ("target", _P)]


class FrameDesc(C.Structure):
    _fields_ = [("mesh", _P), ("shader", C.c_int32), ("enable_shadow", C.c_int32),
                ("uniforms", C.POINTER(Uniforms)), ("diffuse", _P), ("normal", _P), ("shadow_map", _P),
                ("target", _P)]


class PassStats(C.Structure):
    _fields_ = [("faces", C.c_int64), ("clipped_faces", C.c_int64), ("polygons", C.c_int64),
                ("tile_entries", C.c_int64), ("pixels_written", C.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


SERVER_RETURN_DEPTH = 1


class ServerConfig(C.Structure):
    _fields_ = [("lanes", C.c_int32), ("max_batch_frames", C.c_int32), ("max_batch_pixels", C.c_int64),
                ("hbm_budget_bytes", C.c_int64), ("flags", C.c_int32), ("pad_", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("mesh", _P), ("diffuse", _P), ("normal", _P), ("shader", C.c_int32), ("enable_shadow", C.c_int32),
                ("width", C.c_int32), ("height", C.c_int32), ("shadow_width", C.c_int32),
                ("shadow_height", C.c_int32), ("cam_pos", C.c_float * 3), ("cam_target", C.c_float * 3),
                ("light_pos", C.c_float * 3), ("shadow_fp32", C.c_int32), ("user_tag", C.c_uint64)]


class Completion(C.Structure):
    _fields_ = [("id", C.c_uint64), ("user_tag", C.c_uint64), ("status", C.c_int32), ("width", C.c_int32),
                ("height", C.c_int32), ("batch_frames", C.c_int32), ("rgba", C.POINTER(C.c_uint8)),
                ("t_submit_ms", C.c_double), ("t_launch_ms", C.c_double), ("t_done_ms", C.c_double),
                ("depth", C.POINTER(C.c_float))]


class Tuning(C.Structure):
    _fields_ = [("tile_size", C.c_int32), ("raster_wgs_per_cu", C.c_int32), ("parts", C.c_int32),
                ("raster_waves", C.c_int32), ("ms", C.c_double)]

    def as_dict(self):
        return {"tile_size": int(self.tile_size), "raster_wgs_per_cu": int(self.raster_wgs_per_cu),
                "parts": int(self.parts), "raster_waves": int(self.raster_waves), "ms": float(self.ms)}


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64), ("batches", C.c_int64),
                ("retries", C.c_int64), ("inflight_bytes", C.c_int64), ("peak_inflight_bytes", C.c_int64),
                ("busy_ms", C.c_double)]

    def as_dict(self):
        return {k: (float(getattr(self, k)) if k == "busy_ms" else int(getattr(self, k))) for k, _ in self._fields_}


class ObjMesh(C.Structure):
    _fields_ = [("nverts", C.c_int32), ("nfaces", C.c_int32), ("pos", _FP), ("nrm", _FP), ("uv", _FP),
                ("idx", C.POINTER(C.c_int32))]


class ImageS(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("bytespp", C.c_int32), ("pad_", C.c_int32),
                ("bytes", C.POINTER(C.c_uint8))]


DECLS = [
    ("hana_gpu_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("hana_gpu_destroy", C.c_int, [_P]),
    ("hana_gpu_last_error", C.c_char_p, [_P]),
    ("hana_gpu_sync", C.c_int, [_P]),
    ("hana_gpu_set_stats", C.c_int, [_P, C.c_int]),
    ("hana_gpu_last_stats", C.c_int, [_P, C.POINTER(PassStats), C.c_int, C.POINTER(C.c_int)]),
    ("hana_gpu_replay_count", C.c_int, [_P, C.POINTER(C.c_int64)]),
    ("hana_gpu_set_timing", C.c_int, [_P, C.c_int]),
    ("hana_gpu_read_timing", C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    ("hana_mesh_create", C.c_int, [_P, C.c_int32, _FP, _FP, _FP, C.c_int32, C.POINTER(C.c_int32), C.POINTER(_P)]),
    ("hana_mesh_destroy", C.c_int, [_P, _P]),
    ("hana_texture_create", C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_uint8), C.POINTER(_P)]),
    ("hana_texture_destroy", C.c_int, [_P, _P]),
    ("hana_target_create", C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(_P)]),
    ("hana_target_destroy", C.c_int, [_P, _P]),
    ("hana_target_clear", C.c_int, [_P, _P, C.POINTER(C.c_uint8), C.c_float]),
    ("hana_target_upload", C.c_int, [_P, _P, C.POINTER(C.c_uint8), _FP]),
    ("hana_target_read", C.c_int, [_P, _P, C.POINTER(C.c_uint8), _FP]),
    ("hana_target_enable_winner", C.c_int, [_P, _P, C.c_int]),
    ("hana_target_read_winner", C.c_int, [_P, _P, C.POINTER(C.c_uint32)]),
    ("hana_target_present", C.c_int, [_P, _P, C.POINTER(C.c_uint8)]),
    ("hana_draw", C.c_int, [_P, C.POINTER(DrawDesc)]),
    ("hana_render_frame", C.c_int, [_P, C.POINTER(FrameDesc)]),
    ("hana_render_frame_noalloc", C.c_int, [_P, C.POINTER(FrameDesc)]),
    ("hana_gpu_set_tuning", C.c_int, [_P, C.POINTER(Tuning)]),
    ("hana_batch_create", C.c_int, [_P, C.POINTER(_P)]),
    ("hana_batch_set_tuning", C.c_int, [_P, _P, C.POINTER(Tuning)]),
    ("hana_batch_autotune", C.c_int, [_P, _P, C.c_int, C.c_int, C.POINTER(Tuning)]),
    ("hana_batch_get_tuning", C.c_int, [_P, _P, C.POINTER(Tuning)]),
    ("hana_batch_add_frame", C.c_int, [_P, _P, C.POINTER(FrameDesc)]),
    ("hana_batch_finalize", C.c_int, [_P, _P, C.c_int]),
    ("hana_batch_run", C.c_int, [_P, _P]),
    ("hana_batch_destroy", C.c_int, [_P, _P]),
    ("hana_server_create", C.c_int, [_P, C.POINTER(ServerConfig), C.POINTER(_P)]),
    ("hana_server_submit", C.c_int, [_P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    ("hana_server_poll", C.c_int, [_P, C.POINTER(Completion), C.c_int, C.POINTER(C.c_int), C.c_int]),
    ("hana_server_release", C.c_int, [_P, C.c_uint64]),
    ("hana_server_get_stats", C.c_int, [_P, C.POINTER(ServerStats)]),
    ("hana_server_now_ms", C.c_double, [_P]),
    ("hana_server_destroy", C.c_int, [_P]),
    ("hana_build_uniforms", C.c_int, [_FP, _FP, C.c_float, _FP, _FP, _FP, _FP, C.c_int32, C.c_int32, C.c_int32,
                                      C.POINTER(Uniforms)]),
    ("hana_scene_init", C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    ("hana_scene_update_camera", C.c_int, [C.c_void_p, C.c_void_p]),
    ("hana_scene_key", C.c_int, [C.c_void_p, C.c_int32]),
    ("hana_scene_shader", C.c_int, [C.c_void_p]),
    ("hana_scene_uniforms", C.c_int, [C.c_void_p, _FP, _FP, _FP, C.c_int32, C.c_int32, C.POINTER(Uniforms)]),
    ("hana_asset_last_error", C.c_char_p, []),
    ("hana_obj_load", C.c_int, [C.c_char_p, C.POINTER(ObjMesh)]),
    ("hana_obj_free", C.c_int, [C.POINTER(ObjMesh)]),
    ("hana_tga_load", C.c_int, [C.c_char_p, C.c_int, C.POINTER(ImageS)]),
    ("hana_tga_save", C.c_int, [C.c_char_p, C.POINTER(ImageS), C.c_int]),
    ("hana_image_free", C.c_int, [C.POINTER(ImageS)]),
]

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads libhana_gpu.so and declares every entry point (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() (no CPU fallback exists)")
    lib = C.CDLL(path)
    for name, res, args in DECLS:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class HanaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hana error {code}: {msg}")
        self.code = code


def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None:
        assert a.size == n, (a.size, n)
    return a


def _fp(a):
    return a.ctypes.data_as(_FP)


class Motion(C.Structure):
    """hana_motion: Motion (SR/camera.h:11), deltas in window heights."""
    _fields_ = [("orbit", C.c_float * 2), ("pan", C.c_float * 2), ("dolly", C.c_float)]


class SceneStateS(C.Structure):
    _fields_ = [("camera_position", C.c_float * 3), ("camera_target", C.c_float * 3), ("aspect", C.c_float),
                ("light_position", C.c_float * 3), ("shader_index", C.c_int32), ("enable_shadow", C.c_int32)]


KEY_A, KEY_D, KEY_Q, KEY_E, KEY_SPACE = 0, 1, 4, 5, 6


class InteractiveScene:
    """The state SingleModelScene keeps between frames (hana_scene_*): the
    orbit camera, the light, the shader and shadow switches.  frame() gives
    the Scene to render for the current state."""

    def __init__(self, width: int, height: int):
        self.lib = load_library()
        self.width, self.height = int(width), int(height)
        self.s = SceneStateS()
        self._check(self.lib.hana_scene_init(C.byref(self.s), self.width, self.height), "init")

    def _check(self, rc, what):
        if rc < 0:
            raise HanaError(rc, f"hana_scene_{what}")
        return rc

    def move(self, orbit=(0.0, 0.0), pan=(0.0, 0.0), dolly=0.0):
        m = Motion((C.c_float * 2)(*orbit), (C.c_float * 2)(*pan), float(dolly))
        self._check(self.lib.hana_scene_update_camera(C.byref(self.s), C.byref(m)), "update_camera")

    def key(self, k: int):
        self._check(self.lib.hana_scene_key(C.byref(self.s), int(k)), "key")

    @property
    def shader(self) -> int:
        return self._check(self.lib.hana_scene_shader(C.byref(self.s)), "shader")

    @property
    def shadow(self) -> bool:
        return bool(self.s.enable_shadow)

    @property
    def camera_position(self) -> tuple:
        return tuple(self.s.camera_position)

    @property
    def camera_target(self) -> tuple:
        return tuple(self.s.camera_target)

    @property
    def light_position(self) -> tuple:
        return tuple(self.s.light_position)

    def uniforms(self, obj_position=(0.0, 0.0, 0.0), obj_rotation=(0.0, 0.0, 0.0), obj_scale=(1.0, 1.0, 1.0)):
        u = Uniforms()
        arrs = [_f32(x) for x in (obj_position, obj_rotation, obj_scale)]
        self._check(self.lib.hana_scene_uniforms(C.byref(self.s), _fp(arrs[0]), _fp(arrs[1]), _fp(arrs[2]),
                                                 self.width, self.height, C.byref(u)), "uniforms")
        return u

    def frame(self, mesh, diffuse, normal, name="interactive"):
        from . import scenegen as sg
        return sg.Scene(mesh, self.width, self.height, self.shader, self.shadow, self.camera_position,
                        self.light_position, diffuse, normal, cam_target=self.camera_target, name=name)


def build_uniforms(scene, enable_shadow=None) -> Uniforms:
    """DrawModel::draw uniforms computed by the library's host math."""
    lib = load_library()
    u = Uniforms()
    arrs = [_f32(x) for x in (scene.cam_pos, scene.cam_target, scene.light_pos, scene.obj_position,
                              scene.obj_rotation, scene.obj_scale)]
    es = scene.shadow if enable_shadow is None else enable_shadow
    aspect = np.float32(scene.width) / np.float32(scene.height)
    rc = lib.hana_build_uniforms(_fp(arrs[0]), _fp(arrs[1]), float(aspect), _fp(arrs[2]), _fp(arrs[3]), _fp(arrs[4]),
                                 _fp(arrs[5]), scene.width, scene.height, int(es), C.byref(u))
    if rc:
        raise HanaError(rc, "hana_build_uniforms")
    u.shadow_fp32 = int(getattr(scene, "shadow_fp32", False))
    return u


def _asset_check(lib, rc, what):
    if rc:
        msg = lib.hana_asset_last_error()
        raise HanaError(rc, f"{what}: {msg.decode() if msg else ''}")


def load_obj(path):
    """hana_obj_load: a Wavefront OBJ as the renderer's Model reads it, as a
    scenegen Mesh (one vertex per distinct position/uv/fetched normal)."""
    from .scenegen import Mesh
    lib = load_library()
    m = ObjMesh()
    _asset_check(lib, lib.hana_obj_load(os.fsencode(path), C.byref(m)), f"hana_obj_load({path})")
    try:
        V, F = m.nverts, m.nfaces
        pos = np.ctypeslib.as_array(m.pos, (max(V, 1) * 3,))[: 3 * V].reshape(V, 3).copy()
        nrm = np.ctypeslib.as_array(m.nrm, (max(V, 1) * 3,))[: 3 * V].reshape(V, 3).copy()
        uv = np.ctypeslib.as_array(m.uv, (max(V, 1) * 2,))[: 2 * V].reshape(V, 2).copy()
        idx = np.ctypeslib.as_array(m.idx, (max(F, 1) * 3,))[: 3 * F].reshape(F, 3).copy()
    finally:
        lib.hana_obj_free(C.byref(m))
    return Mesh(pos, nrm, uv, idx, fetched_normals=True)


def load_tga(path, flip_vertically=False) -> np.ndarray:
    """hana_tga_load: (H, W) uint8 for 1 byte per texel, else (H, W, bpp), in
    TGAImage memory order (B, G, R, A)."""
    lib = load_library()
    im = ImageS()
    _asset_check(lib, lib.hana_tga_load(os.fsencode(path), int(flip_vertically), C.byref(im)), f"hana_tga_load({path})")
    try:
        n = im.width * im.height * im.bytespp
        a = np.ctypeslib.as_array(im.bytes, (n,)).copy()
    finally:
        shape = (im.height, im.width) if im.bytespp == 1 else (im.height, im.width, im.bytespp)
        lib.hana_image_free(C.byref(im))
    return a.reshape(shape)


def save_tga(path, img: np.ndarray, rle=True) -> None:
    """hana_tga_save: writes a TGAImage-ordered (H, W[, bpp]) uint8 array."""
    lib = load_library()
    a = np.ascontiguousarray(img, np.uint8)
    bpp = 1 if a.ndim == 2 else a.shape[2]
    im = ImageS(a.shape[1], a.shape[0], bpp, 0, a.ctypes.data_as(C.POINTER(C.c_uint8)))
    _asset_check(lib, lib.hana_tga_save(os.fsencode(path), C.byref(im), int(rle)), f"hana_tga_save({path})")


class Context:
    """One HIP stream and the device objects created on it."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = _P()
        rc = self.lib.hana_gpu_create(device, C.byref(h))
        if rc:
            raise HanaError(rc, "hana_gpu_create failed (is a GPU visible?)")
        self.h = h
        self._objs = []

    def check(self, rc, what=""):
        if rc:
            msg = self.lib.hana_gpu_last_error(self.h)
            raise HanaError(rc, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            for kind, o in reversed(self._objs):
                getattr(self.lib, f"hana_{kind}_destroy")(self.h, o)
            self._objs.clear()
            self.lib.hana_gpu_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        self.check(self.lib.hana_gpu_sync(self.h), "sync")

    def set_stats(self, on: bool):
        self.check(self.lib.hana_gpu_set_stats(self.h, int(on)))

    def last_stats(self, max_passes=8):
        arr = (PassStats * max_passes)()
        n = C.c_int()
        self.check(self.lib.hana_gpu_last_stats(self.h, arr, max_passes, C.byref(n)))
        return [arr[i].as_dict() for i in range(n.value)]

    def replay_count(self) -> int:
        '''Waits that grew the workspace and rendered the pending draws again.'''
        n = C.c_int64()
        self.check(self.lib.hana_gpu_replay_count(self.h, C.byref(n)), "replay_count")
        return int(n.value)

    STAGES = ("vertex", "setup", "clip", "scan", "bin", "unused", "raster_g0", "raster_g1")

    def set_tuning(self, tile_size=16, raster_wgs_per_cu=0, parts=4, raster_waves=4):
        '''Tuning of this context's per-call draws (hana_gpu_set_tuning).'''
        t = Tuning(int(tile_size), int(raster_wgs_per_cu), int(parts), int(raster_waves), 0.0)
        self.check(self.lib.hana_gpu_set_tuning(self.h, C.byref(t)), "set_tuning")

    def set_timing(self, on: bool):
        self.check(self.lib.hana_gpu_set_timing(self.h, int(on)))

    def read_timing(self):
        ms = (C.c_double * len(self.STAGES))()
        n = (C.c_int64 * len(self.STAGES))()
        self.check(self.lib.hana_gpu_read_timing(self.h, ms, n))
        return {s: (float(ms[i]), int(n[i])) for i, s in enumerate(self.STAGES)}

    # -- resources --
    def mesh(self, m):
        pos, nrm, uv = _f32(m.pos), _f32(m.nrm), _f32(m.uv)
        idx = np.ascontiguousarray(m.idx, dtype=np.int32)
        h = _P()
        self.check(self.lib.hana_mesh_create(self.h, int(pos.shape[0]), _fp(pos), _fp(nrm), _fp(uv), int(idx.shape[0]),
                                             idx.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(h)), "mesh")
        self._objs.append(("mesh", h))
        return h

    def texture(self, img: np.ndarray):
        img = np.ascontiguousarray(img, dtype=np.uint8)
        hgt, w = img.shape[:2]
        bpp = 1 if img.ndim == 2 else img.shape[2]
        h = _P()
        self.check(self.lib.hana_texture_create(self.h, w, hgt, bpp, img.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                C.byref(h)), "texture")
        self._objs.append(("texture", h))
        return h

    def target(self, w, h, rgba=(0, 0, 0, 255), depth=FLT_MAX, winner=False):
        t = _P()
        self.check(self.lib.hana_target_create(self.h, int(w), int(h), C.byref(t)), "target")
        self._objs.append(("target", t))
        if rgba is not None:
            self.clear(t, rgba, depth)
        if winner:
            self.check(self.lib.hana_target_enable_winner(self.h, t, 1))
        return t

    def clear(self, t, rgba=(0, 0, 0, 255), depth=FLT_MAX):
        b = (C.c_uint8 * 4)(*rgba)
        self.check(self.lib.hana_target_clear(self.h, t, b, float(depth)), "clear")

    def upload(self, t, color: np.ndarray | None, depth: np.ndarray | None):
        cp = np.ascontiguousarray(color, np.uint8).ctypes.data_as(C.POINTER(C.c_uint8)) if color is not None else None
        dp = _fp(np.ascontiguousarray(depth, np.float32)) if depth is not None else None
        self.check(self.lib.hana_target_upload(self.h, t, cp, dp), "upload")

    def read(self, t, w, h):
        color = np.zeros((h, w, 4), np.uint8)
        depth = np.zeros((h, w), np.float32)
        self.check(self.lib.hana_target_read(self.h, t, color.ctypes.data_as(C.POINTER(C.c_uint8)), _fp(depth)), "read")
        return color, depth

    def read_winner(self, t, w, h):
        win = np.zeros((h, w), np.uint32)
        self.check(self.lib.hana_target_read_winner(self.h, t, win.ctypes.data_as(C.POINTER(C.c_uint32))), "winner")
        return win

    def present(self, t, w, h):
        """hana_target_present: the window blit's top-down BGRA8 image."""
        out = np.zeros((h, w, 4), np.uint8)
        self.check(self.lib.hana_target_present(self.h, t, out.ctypes.data_as(C.POINTER(C.c_uint8))), "present")
        return out

    # -- drawing --
    def draw(self, mesh, kind, u: Uniforms, target, diffuse=None, normal=None, shadow_map=None):
        d = DrawDesc(mesh, kind, 0, C.pointer(u), diffuse, normal, shadow_map, target)
        self.check(self.lib.hana_draw(self.h, C.byref(d)), "draw")

    def frame_desc(self, mesh, kind, shadow, u: Uniforms, target, shadow_map=None, diffuse=None, normal=None):
        return FrameDesc(mesh, kind, int(shadow), C.pointer(u), diffuse, normal, shadow_map, target)

    def render_frame(self, mesh, kind, shadow, u: Uniforms, target, shadow_map=None, diffuse=None, normal=None):
        d = self.frame_desc(mesh, kind, shadow, u, target, shadow_map, diffuse, normal)
        self.check(self.lib.hana_render_frame(self.h, C.byref(d)), "render_frame")


class Batch:
    """Many frames rendered by one launch sequence (optionally a hipGraph)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = _P()
        ctx.check(ctx.lib.hana_batch_create(ctx.h, C.byref(h)))
        self.h = h
        self._keep = []

    def add(self, desc: FrameDesc):
        self._keep.append(desc)
        self.ctx.check(self.ctx.lib.hana_batch_add_frame(self.ctx.h, self.h, C.byref(desc)), "batch_add")

    def finalize(self, use_graph=True):
        self.ctx.check(self.ctx.lib.hana_batch_finalize(self.ctx.h, self.h, int(use_graph)), "batch_finalize")

    def set_tuning(self, tile_size=16, raster_wgs_per_cu=0, parts=4, raster_waves=4):
        t = Tuning(int(tile_size), int(raster_wgs_per_cu), int(parts), int(raster_waves), 0.0)
        self.ctx.check(self.ctx.lib.hana_batch_set_tuning(self.ctx.h, self.h, C.byref(t)), "batch_set_tuning")

    def autotune(self, use_graph=True, reps=3) -> dict:
        '''Finalizes with the fastest tuning found (hana_batch_autotune); returns it.'''
        t = Tuning()
        self.ctx.check(self.ctx.lib.hana_batch_autotune(self.ctx.h, self.h, int(use_graph), int(reps), C.byref(t)),
                       "batch_autotune")
        return t.as_dict()

    def tuning(self) -> dict:
        t = Tuning()
        self.ctx.check(self.ctx.lib.hana_batch_get_tuning(self.ctx.h, self.h, C.byref(t)), "batch_get_tuning")
        return t.as_dict()

    def run(self):
        self.ctx.check(self.ctx.lib.hana_batch_run(self.ctx.h, self.h), "batch_run")

    def close(self):
        if self.h:
            self.ctx.lib.hana_batch_destroy(self.ctx.h, self.h)
            self.h = None


class SceneOnDevice:
    """A generated scene's mesh, textures, targets and uniforms on one context."""

    def __init__(self, ctx: Context, scene, winner=False):
        self.ctx, self.scene = ctx, scene
        self.mesh = ctx.mesh(scene.mesh)
        self.diffuse = ctx.texture(scene.diffuse) if scene.diffuse is not None else None
        self.normal = ctx.texture(scene.normal) if scene.normal is not None else None
        self.target = ctx.target(scene.width, scene.height, winner=winner)
        self.shadow_map = None
        if scene.shadow:
            sw, sh = scene.shadow_wh
            self.shadow_map = ctx.target(sw, sh)
        self.u = build_uniforms(scene)

    def desc(self) -> FrameDesc:
        return self.ctx.frame_desc(self.mesh, self.scene.shader, self.scene.shadow, self.u, self.target,
                                   self.shadow_map, self.diffuse, self.normal)

    def render(self):
        self.ctx.render_frame(self.mesh, self.scene.shader, self.scene.shadow, self.u, self.target, self.shadow_map,
                              self.diffuse, self.normal)

    def clear(self):
        self.ctx.clear(self.target)

    def read(self):
        return self.ctx.read(self.target, self.scene.width, self.scene.height)


class Server:
    """Continuous-batching render server (hana_server_*)."""

    def __init__(self, ctx: Context, lanes=2, max_batch_frames=64, max_batch_pixels=64 << 20,
                 hbm_budget_bytes=200 << 30, return_depth=False):
        self.ctx = ctx
        self.lib = ctx.lib
        cfg = ServerConfig(lanes, max_batch_frames, max_batch_pixels, hbm_budget_bytes,
                           SERVER_RETURN_DEPTH if return_depth else 0, 0)
        h = _P()
        ctx.check(self.lib.hana_server_create(ctx.h, C.byref(cfg), C.byref(h)), "server_create")
        self.h = h

    def submit(self, mesh, shader, shadow, width, height, cam_pos, light_pos, diffuse=None, normal=None,
               cam_target=(0.0, 0.0, 0.0), shadow_wh=(0, 0), tag=0, shadow_fp32=False) -> int:
        r = Request(mesh, diffuse, normal, int(shader), int(shadow), int(width), int(height), int(shadow_wh[0]),
                    int(shadow_wh[1]), (C.c_float * 3)(*cam_pos), (C.c_float * 3)(*cam_target),
                    (C.c_float * 3)(*light_pos), int(bool(shadow_fp32)), int(tag))
        rid = C.c_uint64()
        rc = self.lib.hana_server_submit(self.h, C.byref(r), C.byref(rid))
        if rc:
            raise HanaError(rc, "server_submit")
        return int(rid.value)

    def poll(self, max_n=256, wait_ms=10):
        arr = (Completion * max_n)()
        n = C.c_int()
        rc = self.lib.hana_server_poll(self.h, arr, max_n, C.byref(n), int(wait_ms))
        if rc:
            raise HanaError(rc, "server_poll")
        return [arr[i] for i in range(n.value)]

    @staticmethod
    def image(c: Completion) -> np.ndarray:
        n = c.width * c.height * 4
        return np.ctypeslib.as_array(c.rgba, shape=(n,)).reshape(c.height, c.width, 4).copy()

    @staticmethod
    def depth(c: Completion):
        '''The completion's depth buffer (servers made with return_depth), else None.'''
        if not c.depth:
            return None
        return np.ctypeslib.as_array(c.depth, shape=(c.width * c.height,)).reshape(c.height, c.width).copy()

    def release(self, rid: int):
        self.lib.hana_server_release(self.h, C.c_uint64(rid))

    def stats(self) -> dict:
        st = ServerStats()
        self.lib.hana_server_get_stats(self.h, C.byref(st))
        return st.as_dict()

    def now_ms(self) -> float:
        return float(self.lib.hana_server_now_ms(self.h))

    def close(self):
        if self.h:
            self.lib.hana_server_destroy(self.h)
            self.h = None
