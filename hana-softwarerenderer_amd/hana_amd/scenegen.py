"""Seeded synthetic scenes: meshes, textures, cameras and lights.

The reference renders OBJ/TGA assets (SR/model.cpp:6-48, SR/tgaimage.cpp:40-143);
the workloads here are generated in memory instead, deterministically from a
master seed, so that the GPU path and the CPU checker consume the same bytes.

Conventions, shared by every generator:
  * positions, normals, uvs: float32 SoA arrays (V,3), (V,3), (V,2);
    faces: int32 (F,3) indices into them, counter-clockwise seen from outside
    (front-facing under the reference's is_back_facing, SR/graphics.cpp:172-180);
  * normals are fixed points of the reference's Vector3f::normalize
    (SR/vector.h:41-42), so the in-place renormalisation of Model::normal
    (SR/model.cpp:108-111) leaves them unchanged;
  * uvs lie in [0, 1) (u == 1 samples outside the texture, SR/tgaimage.cpp:249);
  * textures are (H, W, 4) uint8 in TGAImage byte order B, G, R, A
    (SR/tgaimage.h:46-53); normal maps encode (n*0.5+0.5)*255 with R = x,
    G = y, B = z as tex_normal decodes them (SR/IShader.h:91-99).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASTER_SEED = 0x48414E41
M64 = (1 << 64) - 1

# shader kinds (same numbering as the C ABI)
GROUND, TOON, TEXTURE, TEXTURE_LIGHT, BLINN, NORMALMAP, SHADOW = range(7)


def splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def config_rng(config_index: int, sub: int = 0) -> np.random.Generator:
    """Per-config stream: splitmix64(master ^ config_index), then sub-streams."""
    s = splitmix64(MASTER_SEED ^ config_index)
    if sub:
        s = splitmix64(s ^ (sub * 0x9E3779B97F4A7C15 & M64))
    return np.random.Generator(np.random.PCG64(s))


# --------------------------------------------------------------------------
# float32 helpers with the reference's operation order
# --------------------------------------------------------------------------

def normalize_ref(v: np.ndarray) -> np.ndarray:
    """Vector3f::normalize: s = 1/sqrt((x*x + y*y) + z*z); v*s (SR/vector.h:41-42)."""
    v = np.asarray(v, dtype=np.float32)
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    n = np.sqrt((x * x + y * y) + z * z)
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.float32(1.0) / n
    return (v * s[:, None]).astype(np.float32)


def _is_fixed(n: np.ndarray) -> np.ndarray:
    return np.all(normalize_ref(n).view(np.uint32) == n.view(np.uint32), axis=1)


def fixed_point_normals(n: np.ndarray) -> np.ndarray:
    """Return unit normals with normalize_ref(n) == n bit for bit.

    Renormalising often alternates between two vectors a few ulps apart; those
    are replaced by the nearest vector (per-component ulp offsets in [-3, 3])
    whose computed norm gives a scale of exactly 1."""
    n = normalize_ref(n)
    n = normalize_ref(n)
    bad = ~_is_fixed(n)
    if not bad.any():
        return n
    rows = np.nonzero(bad)[0]
    base = n[rows].view(np.int32).astype(np.int64)
    found = np.zeros(rows.size, bool)
    out = n[rows].copy()
    offs = sorted(((a, b, c) for a in range(-3, 4) for b in range(-3, 4) for c in range(-3, 4)),
                  key=lambda t: (abs(t[0]) + abs(t[1]) + abs(t[2]), t))
    for o in offs:
        if found.all():
            break
        cand_bits = base.copy()
        # step each component o[k] ulps (sign-magnitude bit patterns)
        sign = np.where(base < 0, -1, 1)
        cand_bits = (base + sign * np.array(o)[None, :]).astype(np.int32)
        cand = cand_bits.view(np.float32)
        ok = _is_fixed(cand) & ~found
        out[ok] = cand[ok]
        found |= ok
    if not found.all():
        raise RuntimeError(f"{int((~found).sum())} normals have no fixed point nearby")
    n[rows] = out
    assert _is_fixed(n).all()
    return n


def clamp_uv(uv: np.ndarray) -> np.ndarray:
    one_minus = np.float32(np.nextafter(np.float32(1.0), np.float32(0.0)))
    return np.clip(uv, 0.0, one_minus).astype(np.float32)


# --------------------------------------------------------------------------
# noise
# --------------------------------------------------------------------------

def _hash_u32(*coords: np.ndarray, seed: int) -> np.ndarray:
    """Integer lattice hash (murmur3-style finaliser over mixed coordinates)."""
    h = np.full(np.broadcast(*coords).shape, np.uint32(seed & 0xFFFFFFFF), dtype=np.uint32)
    primes = (np.uint32(0x9E3779B1), np.uint32(0x85EBCA77), np.uint32(0xC2B2AE3D))
    with np.errstate(over="ignore"):
        for c, p in zip(coords, primes):
            h ^= (c.astype(np.int64) & 0xFFFFFFFF).astype(np.uint32) * p
            h = (h << np.uint32(13)) | (h >> np.uint32(19))
            h = h * np.uint32(5) + np.uint32(0xE6546B64)
        h ^= h >> np.uint32(16)
        h *= np.uint32(0x85EBCA6B)
        h ^= h >> np.uint32(13)
        h *= np.uint32(0xC2B2AE35)
        h ^= h >> np.uint32(16)
    return h


def _smooth(t):
    return t * t * (3.0 - 2.0 * t)


def value_noise_2d(x: np.ndarray, y: np.ndarray, seed: int) -> np.ndarray:
    """Value noise on the grid x (columns, 1-D) by y (rows, 1-D): (len(y), len(x))."""
    xi = np.floor(x).astype(np.int64)
    yi = np.floor(y).astype(np.int64)
    tx = _smooth(x - xi)[None, :]
    ty = _smooth(y - yi)[:, None]
    # hash the lattice points once, then gather rows and columns
    x0, y0 = int(xi.min()), int(yi.min())
    lx = np.arange(x0, int(xi.max()) + 2, dtype=np.int64)
    ly = np.arange(y0, int(yi.max()) + 2, dtype=np.int64)
    lat = _hash_u32(lx[None, :], ly[:, None], seed=seed).astype(np.float64) / 4294967296.0
    ix, iy = xi - x0, yi - y0
    r0, r1 = lat[iy], lat[iy + 1]
    a = r0[:, ix] + (r0[:, ix + 1] - r0[:, ix]) * tx
    b = r1[:, ix] + (r1[:, ix + 1] - r1[:, ix]) * tx
    return a + (b - a) * ty


def value_noise_3d(p: np.ndarray, seed: int) -> np.ndarray:
    i = np.floor(p).astype(np.int64)
    t = _smooth(p - i)
    def g(dx, dy, dz):
        return _hash_u32(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed=seed).astype(np.float64) / 4294967296.0
    c00 = g(0, 0, 0) + (g(1, 0, 0) - g(0, 0, 0)) * t[:, 0]
    c10 = g(0, 1, 0) + (g(1, 1, 0) - g(0, 1, 0)) * t[:, 0]
    c01 = g(0, 0, 1) + (g(1, 0, 1) - g(0, 0, 1)) * t[:, 0]
    c11 = g(0, 1, 1) + (g(1, 1, 1) - g(0, 1, 1)) * t[:, 0]
    c0 = c00 + (c10 - c00) * t[:, 1]
    c1 = c01 + (c11 - c01) * t[:, 1]
    return c0 + (c1 - c0) * t[:, 2]


def fbm_2d(w: int, h: int, seed: int, base_cells: float = 8.0, octaves: int = 4,
           persistence: float = 0.5) -> np.ndarray:
    """4-octave value noise over a w x h image, in [0, 1)."""
    ys, xs = np.arange(h, dtype=np.float64), np.arange(w, dtype=np.float64)
    acc = np.zeros((h, w))
    amp, tot = 1.0, 0.0
    for o in range(octaves):
        f = base_cells * (2 ** o)
        acc += amp * value_noise_2d(xs * f / w, ys * f / h, seed + 7919 * o)
        tot += amp
        amp *= persistence
    return acc / tot


def fbm_3d(p: np.ndarray, seed: int, freq: float = 4.0, octaves: int = 4,
           persistence: float = 0.5) -> np.ndarray:
    acc = np.zeros(p.shape[0])
    amp, tot = 1.0, 0.0
    for o in range(octaves):
        acc += amp * value_noise_3d(p * (freq * 2 ** o), seed + 104729 * o)
        tot += amp
        amp *= persistence
    return acc / tot


def diffuse_texture(size: int, seed: int) -> np.ndarray:
    """Value-noise albedo, RGBA8 in TGA byte order (B, G, R, A)."""
    n0 = fbm_2d(size, size, seed)
    n1 = fbm_2d(size, size, seed + 1, base_cells=16.0)
    r = 0.35 + 0.6 * n0
    g = 0.25 + 0.5 * (0.5 * n0 + 0.5 * n1)
    b = 0.2 + 0.6 * n1
    img = np.empty((size, size, 4), dtype=np.uint8)
    img[..., 0] = np.clip(b * 255, 0, 255).astype(np.uint8)
    img[..., 1] = np.clip(g * 255, 0, 255).astype(np.uint8)
    img[..., 2] = np.clip(r * 255, 0, 255).astype(np.uint8)
    img[..., 3] = 255
    return img


def normal_texture(size: int, seed: int, strength: float = 6.0) -> np.ndarray:
    """Tangent-space normal map from the gradient of a value-noise height."""
    hgt = fbm_2d(size, size, seed, base_cells=12.0)
    dx = (np.roll(hgt, -1, axis=1) - np.roll(hgt, 1, axis=1)) * 0.5 * strength
    dy = (np.roll(hgt, -1, axis=0) - np.roll(hgt, 1, axis=0)) * 0.5 * strength
    n = np.stack([-dx, -dy, np.ones_like(dx)], axis=-1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    enc = np.clip((n * 0.5 + 0.5) * 255.0, 0, 255).astype(np.uint8)
    img = np.empty((size, size, 4), dtype=np.uint8)
    img[..., 0] = enc[..., 2]   # B = z
    img[..., 1] = enc[..., 1]   # G = y
    img[..., 2] = enc[..., 0]   # R = x
    img[..., 3] = 255
    return img


# --------------------------------------------------------------------------
# meshes
# --------------------------------------------------------------------------

@dataclass
class Mesh:
    pos: np.ndarray      # (V,3) float32
    nrm: np.ndarray      # (V,3) float32
    uv: np.ndarray       # (V,2) float32
    idx: np.ndarray      # (F,3) int32
    # normals already are what Model::normal returns at each corner (an OBJ
    # read by hana_obj_load); otherwise they are stored vectors that every
    # fetch renormalises (generated meshes hold fixed points, so both agree)
    fetched_normals: bool = False

    @property
    def nverts(self) -> int:
        return int(self.pos.shape[0])

    @property
    def nfaces(self) -> int:
        return int(self.idx.shape[0])


def _orient_outward(pos: np.ndarray, idx: np.ndarray, center: np.ndarray | None = None) -> np.ndarray:
    """Make faces counter-clockwise seen from outside (normal points away from center)."""
    a, b, c = pos[idx[:, 0]].astype(np.float64), pos[idx[:, 1]].astype(np.float64), pos[idx[:, 2]].astype(np.float64)
    fn = np.cross(b - a, c - a)
    ctr = (a + b + c) / 3.0 - (0.0 if center is None else center)
    flip = np.einsum("ij,ij->i", fn, ctr) < 0
    idx = idx.copy()
    idx[flip] = idx[flip][:, [0, 2, 1]]
    return idx


def vertex_normals(pos: np.ndarray, idx: np.ndarray) -> np.ndarray:
    a, b, c = pos[idx[:, 0]].astype(np.float64), pos[idx[:, 1]].astype(np.float64), pos[idx[:, 2]].astype(np.float64)
    fn = np.cross(b - a, c - a)
    n = np.zeros((pos.shape[0], 3))
    for k in range(3):
        np.add.at(n, idx[:, k], fn)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    unused = ln[:, 0] == 0
    n[unused] = (0.0, 0.0, 1.0)      # vertices no face references
    ln[unused] = 1.0
    return (n / ln).astype(np.float32)


def uv_sphere(segments: int = 64, rings: int = 32, radius: float = 1.0) -> Mesh:
    """Latitude/longitude sphere with a duplicated seam column; poles are
    single triangles, so F = 2*segments*(rings-1) and V = (segments+1)*(rings+1)."""
    i = np.arange(rings + 1, dtype=np.float64)[:, None]
    j = np.arange(segments + 1, dtype=np.float64)[None, :]
    theta = np.pi * i / rings
    phi = 2.0 * np.pi * j / segments
    x = np.sin(theta) * np.cos(phi)
    y = np.cos(theta) * np.ones_like(phi)
    z = -np.sin(theta) * np.sin(phi)
    unit = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    pos = (unit * radius).astype(np.float32)
    u = (j / segments) * np.ones_like(theta)
    v = (1.0 - i / rings) * np.ones_like(phi)
    uv = clamp_uv(np.stack([u, v], axis=-1).reshape(-1, 2).astype(np.float32))
    vid = lambda r, s: r * (segments + 1) + s
    faces = []
    for r in range(rings):
        for s in range(segments):
            a, b, c, d = vid(r, s), vid(r, s + 1), vid(r + 1, s), vid(r + 1, s + 1)
            if r != 0:
                faces.append((a, c, b))
            if r != rings - 1:
                faces.append((b, c, d))
    idx = _orient_outward(pos, np.asarray(faces, dtype=np.int32))
    nrm = fixed_point_normals(unit.astype(np.float32))
    return Mesh(pos, nrm, uv, idx.astype(np.int32))


def icosphere(levels: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere subdivided `levels` times: (V,3) float64, (F,3) int64."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(levels):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a, b = uniq // nv, uniq % nv
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = (inv + nv).reshape(3, -1).T   # midpoint index of edges (01, 12, 20) per face
        v = np.concatenate([v, mid], axis=0)
        f0, f1, f2 = f[:, 0], f[:, 1], f[:, 2]
        m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
        # the four children of face i are faces 4i..4i+3: submission order
        # stays spatially coherent at every level
        f = np.stack([
            np.stack([f0, m01, m20], 1), np.stack([f1, m12, m01], 1),
            np.stack([f2, m20, m12], 1), np.stack([m01, m12, m20], 1)], axis=1).reshape(-1, 3)
    return reorder_vertices(v, f)


def reorder_vertices(v: np.ndarray, f: np.ndarray):
    """Renumbers vertices in order of first use by the faces."""
    flat = f.ravel()
    uniq, first = np.unique(flat, return_index=True)
    order = uniq[np.argsort(first)]
    remap = np.empty(v.shape[0], np.int64)
    remap[order] = np.arange(order.size)
    return v[order], remap[f]


def spherical_uv(unit: np.ndarray) -> np.ndarray:
    u = 0.5 + np.arctan2(unit[:, 2], unit[:, 0]) / (2 * np.pi)
    vv = 0.5 + np.arcsin(np.clip(unit[:, 1], -1, 1)) / np.pi
    return clamp_uv(np.stack([u, vv], axis=1).astype(np.float32))


def displaced_icosphere(levels: int, amp: float, seed: int, freq: float = 3.0) -> Mesh:
    unit, f = icosphere(levels)
    r = 1.0 + amp * (2.0 * fbm_3d(unit, seed, freq=freq) - 1.0)
    pos = (unit * r[:, None]).astype(np.float32)
    idx = _orient_outward(pos, f.astype(np.int32))
    nrm = fixed_point_normals(vertex_normals(pos, idx))
    return Mesh(pos, nrm, spherical_uv(unit), idx.astype(np.int32))


def displaced_uv_sphere(segments: int, rings: int, amp: float, seed: int) -> Mesh:
    m = uv_sphere(segments, rings)
    unit = m.pos.astype(np.float64)
    r = 1.0 + amp * (2.0 * fbm_3d(unit, seed) - 1.0)
    pos = (unit * r[:, None]).astype(np.float32)
    nrm = fixed_point_normals(vertex_normals(pos, m.idx))
    return Mesh(pos, nrm, m.uv, m.idx)


def box(n: int) -> Mesh:
    """Unit cube [-1,1]^3, each face an n x n grid with its own vertices."""
    g = np.linspace(-1.0, 1.0, n + 1)
    a, b = np.meshgrid(g, g, indexing="ij")
    a, b = a.ravel(), b.ravel()
    one = np.ones_like(a)
    faces_pos, faces_nrm, faces_uv, faces_idx = [], [], [], []
    base = 0
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            p = np.zeros((a.size, 3))
            p[:, axis] = sgn * one
            p[:, (axis + 1) % 3] = a
            p[:, (axis + 2) % 3] = b
            nn = np.zeros((a.size, 3))
            nn[:, axis] = sgn
            uvs = np.stack([(a + 1) / 2, (b + 1) / 2], axis=1)
            q = np.arange((n + 1) * (n + 1)).reshape(n + 1, n + 1)
            t1 = np.stack([q[:-1, :-1].ravel(), q[1:, :-1].ravel(), q[:-1, 1:].ravel()], 1)
            t2 = np.stack([q[1:, :-1].ravel(), q[1:, 1:].ravel(), q[:-1, 1:].ravel()], 1)
            faces_pos.append(p)
            faces_nrm.append(nn)
            faces_uv.append(uvs)
            faces_idx.append(np.stack([t1, t2], 1).reshape(-1, 3) + base)
            base += a.size
    pos = np.concatenate(faces_pos).astype(np.float32)
    idx = _orient_outward(pos, np.concatenate(faces_idx).astype(np.int32))
    return Mesh(pos, fixed_point_normals(np.concatenate(faces_nrm).astype(np.float32)),
                clamp_uv(np.concatenate(faces_uv).astype(np.float32)), idx.astype(np.int32))


def rotation_matrix(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def place(mesh: Mesh, rot: np.ndarray, scale: float, offset: np.ndarray) -> Mesh:
    pos = (mesh.pos.astype(np.float64) @ rot.T) * scale + offset
    nrm = mesh.nrm.astype(np.float64) @ rot.T
    return Mesh(pos.astype(np.float32), fixed_point_normals(nrm.astype(np.float32)), mesh.uv, mesh.idx)


def merge(meshes: list[Mesh]) -> Mesh:
    off = np.cumsum([0] + [m.nverts for m in meshes[:-1]])
    return Mesh(np.concatenate([m.pos for m in meshes]), np.concatenate([m.nrm for m in meshes]),
                np.concatenate([m.uv for m in meshes]),
                np.concatenate([m.idx + o for m, o in zip(meshes, off)]).astype(np.int32))


def sphere_with_tris(target: int, amp: float, seed: int) -> Mesh:
    """UV sphere whose face count 2*s*(r-1) is close to target (s = 2r)."""
    r = max(3, int(round((1 + math.sqrt(1 + 4 * target)) / 4)))
    s = 2 * r
    return displaced_uv_sphere(s, r, amp, seed) if amp > 0 else uv_sphere(s, r)


def box_with_tris(target: int) -> Mesh:
    return box(max(1, int(round(math.sqrt(target / 12.0)))))


# --------------------------------------------------------------------------
# scenes
# --------------------------------------------------------------------------

@dataclass
class Scene:
    """One frame: DrawModel::draw over a mesh (SR/scene.h:53-99)."""
    mesh: Mesh
    width: int
    height: int
    shader: int
    shadow: bool
    cam_pos: tuple
    light_pos: tuple
    diffuse: np.ndarray
    normal: np.ndarray
    shadow_size: tuple | None = None    # (w, h); None = frame-sized, as the reference
    cam_target: tuple = (0.0, 0.0, 0.0)
    obj_position: tuple = (0.0, 0.0, 0.0)
    obj_rotation: tuple = (0.0, 0.0, 0.0)
    obj_scale: tuple = (1.0, 1.0, 1.0)
    name: str = ""
    meta: dict = field(default_factory=dict)
    shadow_fp32: bool = False           # the fp32 shadow-map extension (hana_uniforms::shadow_fp32)

    @property
    def aspect(self) -> float:
        return float(np.float32(self.width) / np.float32(self.height))

    @property
    def shadow_wh(self) -> tuple:
        return self.shadow_size if self.shadow_size else (self.width, self.height)


def random_light(rng: np.random.Generator, radius: float = 2.0 * math.sqrt(3.0), front: bool = False) -> tuple:
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    if d[1] < 0.2:             # keep the light above the horizon so that lit areas exist
        d[1] = abs(d[1]) + 0.2
        d /= np.linalg.norm(d)
    if front and d[2] < 0.3:   # on the camera's side of layers facing +z
        d[2] = abs(d[2]) + 0.3
        d /= np.linalg.norm(d)
    return tuple(float(np.float32(x)) for x in d * radius)


def orbit_camera(rng: np.random.Generator, dist_lo: float, dist_hi: float) -> tuple:
    az = rng.uniform(0, 2 * np.pi)
    el = rng.uniform(-0.6, 0.9)
    dist = rng.uniform(dist_lo, dist_hi)
    p = np.array([math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az)]) * dist
    return tuple(float(np.float32(x)) for x in p)


_TEX_CACHE: dict = {}


def texture_pair(diffuse_size: int, normal_size: int, seed: int):
    key = (diffuse_size, normal_size, seed)
    if key not in _TEX_CACHE:
        _TEX_CACHE[key] = (diffuse_texture(diffuse_size, seed), normal_texture(normal_size, seed + 101))
    return _TEX_CACHE[key]


def config1_scene(shader: int = NORMALMAP, width: int = 800, height: int = 600) -> Scene:
    """Plain method, small: UV sphere 64x32, 1024^2 diffuse + 512^2 normal map, 800x600, no shadow."""
    rng = config_rng(0)
    mesh = uv_sphere(64, 32)
    d, n = texture_pair(1024, 512, 11)
    return Scene(mesh, width, height, shader, False, (0.0, 0.0, 2.0), random_light(rng), d, n, name="C1")


def config2_scene(levels: int = 8, width: int = 1920, height: int = 1080, shadow_size: int = 2048,
                  tex_size: int = 2048) -> Scene:
    """Single large frame: icosphere 8x (1.31M tris), displacement amp 0.05,
    2048^2 textures, NormalMap + 2048^2 shadow pre-pass, 1920x1080."""
    rng = config_rng(1)
    mesh = displaced_icosphere(levels, 0.05, seed=int(rng.integers(1 << 30)))
    d, n = texture_pair(tex_size, tex_size, 21)
    return Scene(mesh, width, height, NORMALMAP, True, (0.0, 0.0, 2.0), random_light(rng), d, n,
                 shadow_size=(shadow_size, shadow_size), name="C2")


def log_uniform(rng, lo, hi):
    return float(math.exp(rng.uniform(math.log(lo), math.log(hi))))


def config3_scenes(count: int = 256, tri_lo: int = 1000, tri_hi: int = 200_000,
                   px_lo: int = 128 * 128, px_hi: int = 1024 * 1024) -> list[Scene]:
    """Batched: independent scenes of 1-8 displaced spheres / boxes, log-uniform
    triangle totals and pixel counts, random aspect in [1/2, 2], shader in
    {Blinn, NormalMap} x shadow in {off, on}."""
    rng = config_rng(2)
    texs = [texture_pair(512, 512, 31 + k) for k in range(4)]
    scenes = []
    for s in range(count):
        total = int(log_uniform(rng, tri_lo, tri_hi))
        k = int(rng.integers(1, 9))
        parts = rng.dirichlet(np.ones(k)) * total
        meshes = []
        for p in parts:
            t = max(64, int(p))
            if rng.uniform() < 0.6:
                m = sphere_with_tris(t, 0.05, seed=int(rng.integers(1 << 30)))
            else:
                m = box_with_tris(t)
            off = rng.uniform(-0.55, 0.55, size=3)
            meshes.append(place(m, rotation_matrix(rng), float(rng.uniform(0.18, 0.4)), off))
        mesh = merge(meshes)
        npx = log_uniform(rng, px_lo, px_hi)
        aspect = log_uniform(rng, 0.5, 2.0)
        w = max(16, int(round(math.sqrt(npx * aspect))))
        h = max(16, int(round(math.sqrt(npx / aspect))))
        shader = BLINN if rng.uniform() < 0.5 else NORMALMAP
        shadow = bool(rng.uniform() < 0.5)
        d, n = texs[int(rng.integers(len(texs)))]
        scenes.append(Scene(mesh, w, h, shader, shadow, orbit_camera(rng, 2.0, 2.8), random_light(rng),
                            d, n, name=f"C3/{s}"))
    return scenes


def config5_scene(layers: int = 8, grid: int = 1119, width: int = 3840, height: int = 2160,
                  shadow_size: int = 4096, tex_size: int = 2048) -> Scene:
    """Micro-triangle stress: `layers` jittered grid layers of grid^2 vertices
    facing the camera and filling the view (triangles of 1-4 px at 3840x2160).
    Layers come in pairs with identical positions (exact depth ties broken by
    submission order) and differently offset uvs; the pairs sit at random
    depths.  NormalMap + 4096^2 shadow pre-pass."""
    rng = config_rng(4)
    cam_z = 2.0
    aspect = width / height
    tan_half = math.tan(math.radians(30.0))
    meshes = []
    depths = rng.uniform(-0.6, 0.6, size=(layers + 1) // 2)
    g = np.arange(grid, dtype=np.float64) / (grid - 1)
    gx, gy = np.meshgrid(g, g, indexing="xy")
    q = np.arange(grid * grid).reshape(grid, grid)
    t1 = np.stack([q[:-1, :-1].ravel(), q[:-1, 1:].ravel(), q[1:, :-1].ravel()], 1)
    t2 = np.stack([q[:-1, 1:].ravel(), q[1:, 1:].ravel(), q[1:, :-1].ravel()], 1)
    idx = np.stack([t1, t2], 1).reshape(-1, 3).astype(np.int32)
    for pair in range(len(depths)):
        z = float(depths[pair])
        dist = cam_z - z
        hh = dist * tan_half * 1.02
        hw = hh * aspect
        jit = rng.uniform(-0.3, 0.3, size=(grid, grid, 2)) / (grid - 1)
        jit[0, :, :] = 0; jit[-1, :, :] = 0; jit[:, 0, :] = 0; jit[:, -1, :] = 0
        x = (gx + jit[..., 0]) * 2 * hw - hw
        y = (gy + jit[..., 1]) * 2 * hh - hh
        pos = np.stack([x.ravel(), y.ravel(), np.full(x.size, z)], 1).astype(np.float32)
        nrm = np.tile(np.array([[0.0, 0.0, 1.0]], dtype=np.float32), (pos.shape[0], 1))
        for rep in range(2):
            if len(meshes) == layers:
                break
            shift = rng.uniform(0, 0.5, size=2)
            uv = clamp_uv(np.stack([(gx.ravel() * 0.5 + shift[0]), (gy.ravel() * 0.5 + shift[1])], 1).astype(np.float32))
            meshes.append(Mesh(pos, nrm, uv, idx))
    # fix winding with respect to the camera (+z)
    mesh = merge(meshes)
    a, b, c = (mesh.pos[mesh.idx[:, k]].astype(np.float64) for k in range(3))
    fz = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    flip = fz < 0
    mesh.idx[flip] = mesh.idx[flip][:, [0, 2, 1]]
    d, n = texture_pair(tex_size, tex_size, 51)
    # the light faces the layers too, so that the shadow pre-pass rasterises them
    return Scene(mesh, width, height, NORMALMAP, True, (0.0, 0.0, cam_z), random_light(rng, front=True), d, n,
                 shadow_size=(shadow_size, shadow_size), name="C5")


def small_scene(kind: int = NORMALMAP, shadow: bool = False, width: int = 160, height: int = 120,
                seed: int = 0, segments: int = 24, rings: int = 12, cam=(0.0, 0.0, 2.0),
                shadow_size=None, tex: int = 64) -> Scene:
    """Small seeded frames for parity tests."""
    rng = np.random.Generator(np.random.PCG64(splitmix64(MASTER_SEED ^ (1000 + seed))))
    mesh = displaced_uv_sphere(segments, rings, 0.08, seed=int(rng.integers(1 << 30)))
    d, n = texture_pair(tex, tex, 900 + seed)
    return Scene(mesh, width, height, kind, shadow, cam, random_light(rng), d, n,
                 shadow_size=shadow_size, name=f"small/{seed}")


@dataclass
class ServeWorkload:
    """Config 4: resident meshes and a Poisson stream of render requests."""
    meshes: list
    textures: list
    requests: list          # dicts: t_ms, mesh, shader, shadow, width, height, cam_pos, light_pos, tex


def config4_workload(n_meshes: int = 64, tri_lo: int = 5000, tri_hi: int = 500_000, rate_per_s: float = 2000.0,
                     seconds: float = 60.0, px_lo: int = 256 * 256, px_hi: int = 1920 * 1080,
                     max_requests: int | None = None) -> ServeWorkload:
    """64 device-resident meshes (log-uniform 5k-500k triangles), Poisson
    arrivals at rate_per_s; each request a random orbit camera and light over
    one mesh at a log-uniform pixel count in [256^2, 1920x1080] with random
    aspect, Blinn or NormalMap, shadow on or off."""
    rng = config_rng(3)
    texs = [texture_pair(1024, 1024, 61 + k) for k in range(4)]
    meshes = []
    for i in range(n_meshes):
        t = int(log_uniform(rng, tri_lo, tri_hi))
        if rng.uniform() < 0.75:
            meshes.append(sphere_with_tris(t, 0.05, seed=int(rng.integers(1 << 30))))
        else:
            m = box_with_tris(t)
            meshes.append(place(m, rotation_matrix(rng), 0.8, np.zeros(3)))
    reqs = []
    t = 0.0
    limit = max_requests if max_requests is not None else 1 << 62
    while len(reqs) < limit:
        t += rng.exponential(1000.0 / rate_per_s)
        if t > seconds * 1000.0:
            break
        npx = log_uniform(rng, px_lo, px_hi)
        aspect = log_uniform(rng, 0.5, 2.0)
        w = max(16, min(4096, int(round(math.sqrt(npx * aspect)))))
        h = max(16, min(4096, int(round(math.sqrt(npx / aspect)))))
        reqs.append({"t_ms": t, "mesh": int(rng.integers(n_meshes)),
                     "shader": BLINN if rng.uniform() < 0.5 else NORMALMAP, "shadow": bool(rng.uniform() < 0.5),
                     "width": w, "height": h, "cam_pos": orbit_camera(rng, 2.2, 3.2), "light_pos": random_light(rng),
                     "tex": int(rng.integers(len(texs)))})
    return ServeWorkload(meshes, texs, reqs)


def request_scene(wl: ServeWorkload, req: dict) -> Scene:
    """The DrawModel::draw frame a request renders."""
    d, n = wl.textures[req["tex"]]
    return Scene(wl.meshes[req["mesh"]], req["width"], req["height"], req["shader"], req["shadow"],
                 tuple(req["cam_pos"]), tuple(req["light_pos"]), d, n, name=f"C4/{req['t_ms']:.3f}")
