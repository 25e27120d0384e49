"""Scenes for edge cases of the depth test, shared by the oracle and GPU tests."""
import numpy as np

from hana_amd import scenegen as sg


def _quad(x0, y0, x1, y1, pz):
    p = np.array([[x0, y0, pz], [x1, y0, pz], [x1, y1, pz], [x0, y1, pz]], np.float32)
    return p, np.array([[0, 1, 2], [0, 2, 3]], np.int32)


# Clip transform of the NaN-depth scene: x = s*px, y = s*py, z = s*pz - s,
# w = s*pz with s = 1e10, so ndc = (px/pz, py/pz, 1 - 1/pz).  A vertex at
# pz = 1e30 overflows z and w to +inf: it passes the trivial accept
# (|x|, |y|, |z| <= w), lands on the centre of the screen and its depth
# (inf/inf + 1) * 0.5 is NaN, so every fragment of its triangles has NaN depth.
NAN_SCENE_VP = np.zeros((4, 4), np.float32)
NAN_SCENE_VP[0, 0] = NAN_SCENE_VP[1, 1] = NAN_SCENE_VP[2, 2] = NAN_SCENE_VP[3, 2] = np.float32(1e10)
NAN_SCENE_VP[2, 3] = -np.float32(1e10)


def nan_depth_scene(width=96, height=72, seed=3, shift=0.0):
    """A near quad, then a triangle of NaN depth over it, then two farther
    quads: the NaN fragments replace the near quad's depth (NaN > d is false),
    and every fragment after the last NaN one at a pixel passes.  `shift`
    moves the NaN triangle's far corners, for a second draw into the same
    target."""
    parts = [_quad(-0.8, -0.8, 0.8, 0.8, 2.0),
             (np.array([[0.3, 0.2, 1e30], [-1.6 + shift, -1.4, 2.0], [1.5, -1.2 + shift, 2.0]], np.float32),
              np.array([[0, 1, 2]], np.int32)),
             _quad(-0.6, -1.0, 1.0, 0.5, 3.0),
             _quad(-1.0, -0.2, 0.4, 1.0, 2.5)]
    pos, idx, n = [], [], 0
    for p, i in parts:
        pos.append(p)
        idx.append(i + n)
        n += len(p)
    pos = np.concatenate(pos)
    rng = np.random.default_rng(seed)
    uv = rng.uniform(0, 1, (len(pos), 2)).astype(np.float32)
    nrm = np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (len(pos), 1))
    mesh = sg.Mesh(pos, nrm, uv, np.concatenate(idx))
    base = sg.small_scene(sg.TEXTURE, width=width, height=height, seed=seed)
    return sg.Scene(mesh, width, height, sg.TEXTURE, False, (0.0, 0.0, 2.0), base.light_pos, base.diffuse,
                    base.normal, name=f"nan{shift}")


def with_camera_vp(u, m=NAN_SCENE_VP):
    """The uniforms u with camera_vp replaced by m (row-major)."""
    for i, v in enumerate(np.asarray(m, np.float32).ravel()):
        u.camera_vp[i] = v
    return u


def sliver_scene(n=800_000, width=2048, height=2048, seed=11):
    """n slivers 1 pixel wide and 32 tall, each straddling a tile column
    boundary and spanning three tile rows: six (triangle, tile) pairs per
    triangle, more than the initial bin capacity of four per face, with random
    depths.  With camera_vp the identity, positions are NDC."""
    rng = np.random.default_rng(seed)
    tx = rng.integers(1, width // 16, n)
    ty = rng.integers(0, height // 16 - 2, n)
    sx = tx * 16 - 0.5 + rng.uniform(-0.2, 0.2, n)
    sy = ty * 16 + 8 + rng.uniform(-2.0, 2.0, n)
    corners = np.stack([np.stack([sx, sy], 1), np.stack([sx + 1.0, sy + 2.0], 1), np.stack([sx + 0.5, sy + 32.0], 1)], 1)
    ndc = corners / np.array([width, height]) * 2 - 1
    z = np.repeat(rng.uniform(-0.9, 0.9, (n, 1, 1)), 3, 1)
    pos = np.concatenate([ndc, z], 2).reshape(-1, 3).astype(np.float32)
    uv = rng.uniform(0, 1, (3 * n, 2)).astype(np.float32)
    nrm = np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (3 * n, 1))
    mesh = sg.Mesh(pos, nrm, uv, np.arange(3 * n, dtype=np.int32).reshape(n, 3))
    base = sg.small_scene(sg.TEXTURE, width=256, height=256, seed=seed)
    return sg.Scene(mesh, width, height, sg.TEXTURE, False, (0.0, 0.0, 2.0), base.light_pos, base.diffuse,
                    base.normal, name="slivers")


IDENTITY_VP = np.eye(4, dtype=np.float32)


def empty_mesh():
    """A mesh with no vertices and no faces."""
    return sg.Mesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), np.zeros((0, 2), np.float32),
                   np.zeros((0, 3), np.int32))


def faceless(mesh):
    """The vertices of mesh with no faces."""
    return sg.Mesh(mesh.pos, mesh.nrm, mesh.uv, np.zeros((0, 3), np.int32))


def tiny_input_scene(n=4000, width=256, height=192, seed=21):
    """Triangles whose screen-space set-up has tiny inputs, beside ordinary
    ones: a corner within 2^-8 px of the left or top edge of the target, edges
    a few ulps from vertical or horizontal (a screen component of 2^-16 px or
    so), and corners exactly at 2^-8 px from the edge and one ulp inside it.
    The raster's coverage takes its shared-reciprocal shortcut only for
    triangles whose inputs are each zero or at least 2^-8 px, and divides for
    the others, so both paths and their boundary are drawn.  Sizes run from a
    pixel to a third of the target, so the per-lane, shared and wave-wide
    sweeps all see them.  With camera_vp the identity, positions are NDC."""
    rng = np.random.default_rng(seed)
    f32 = np.float32
    W, H = f32(width), f32(height)

    def ndc_x(sx):
        return f32(f32(sx) / W * f32(2) - f32(1))

    def ndc_y(sy):
        return f32(f32(sy) / H * f32(2) - f32(1))

    tris = []
    for i in range(n):
        kind = i % 6
        size = float(np.exp(rng.uniform(np.log(1.0), np.log(min(width, height) / 3))))
        ax, ay = rng.uniform(0, width - 1), rng.uniform(0, height - 1)
        a = [ndc_x(ax), ndc_y(ay)]
        b = [ndc_x(ax + rng.uniform(-1, 1) * size), ndc_y(ay + rng.uniform(-1, 1) * size)]
        c = [ndc_x(ax + rng.uniform(-1, 1) * size), ndc_y(ay + rng.uniform(-1, 1) * size)]
        if kind == 0:      # a corner within 2^-8 px of the left edge: ndc -1 + k 2^-24
            a[0] = f32(-1) + f32(int(rng.integers(1, 400))) * f32(2.0 ** -24)
            b[0] = ndc_x(rng.uniform(0.5, size + 1))
            c[0] = ndc_x(rng.uniform(0.5, size + 1))
        elif kind == 1:    # the same at the top edge
            a[1] = f32(-1) + f32(int(rng.integers(1, 400))) * f32(2.0 ** -24)
            b[1] = ndc_y(rng.uniform(0.5, size + 1))
            c[1] = ndc_y(rng.uniform(0.5, size + 1))
        elif kind == 2:    # edge a-b a few ulps from vertical
            b[0] = a[0]
            for _ in range(int(rng.integers(1, 4))):
                b[0] = np.nextafter(b[0], f32(np.inf) if rng.uniform() < 0.5 else f32(-np.inf))
        elif kind == 3:    # edge a-c a few ulps from horizontal
            c[1] = a[1]
            for _ in range(int(rng.integers(1, 4))):
                c[1] = np.nextafter(c[1], f32(np.inf) if rng.uniform() < 0.5 else f32(-np.inf))
        elif kind == 4:    # a corner at 2^-8 px from the left edge, or one ulp inside
            a[0] = f32(-1) + f32(2.0 ** -15) - (f32(2.0 ** -24) if rng.uniform() < 0.5 else f32(0))
            b[0] = ndc_x(rng.uniform(0.5, size + 1))
            c[0] = ndc_x(rng.uniform(0.5, size + 1))
        z = rng.uniform(-0.9, 0.9, 3).astype(f32)
        p = np.array([[a[0], a[1], z[0]], [b[0], b[1], z[1]], [c[0], c[1], z[2]]], f32)
        area = (float(p[1, 0]) - float(p[0, 0])) * (float(p[2, 1]) - float(p[0, 1])) - \
               (float(p[2, 0]) - float(p[0, 0])) * (float(p[1, 1]) - float(p[0, 1]))
        if area < 0:
            p = p[[0, 2, 1]]
        tris.append(p)
    pos = np.concatenate(tris).astype(f32)
    uv = rng.uniform(0, 1, (3 * n, 2)).astype(f32)
    nrm = np.tile(np.array([[0.0, 0.0, 1.0]], f32), (3 * n, 1))
    mesh = sg.Mesh(pos, nrm, uv, np.arange(3 * n, dtype=np.int32).reshape(n, 3))
    base = sg.small_scene(sg.TEXTURE, width=64, height=64, seed=seed)
    return sg.Scene(mesh, width, height, sg.TEXTURE, False, (0.0, 0.0, 2.0), base.light_pos, base.diffuse,
                    base.normal, name="tiny inputs")


def screen_inputs(scene):
    """Per face, the six screen-space set-up inputs (a's position, both edges)
    as the raster forms them from NDC corners, in float32."""
    f32 = np.float32
    p = scene.mesh.pos[scene.mesh.idx]           # (F, 3, 3) NDC with camera_vp the identity
    W, H = f32(scene.width), f32(scene.height)
    sx = (p[..., 0] + f32(1)) * f32(0.5) * W
    sy = (p[..., 1] + f32(1)) * f32(0.5) * H
    return np.stack([sx[:, 0], sy[:, 0], sx[:, 2] - sx[:, 0], sx[:, 1] - sx[:, 0],
                     sy[:, 2] - sy[:, 0], sy[:, 1] - sy[:, 0]], 1).astype(f32)
