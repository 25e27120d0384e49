"""Parity helpers: the GPU result against the oracle's.  The bar: depth
bit-exact, colour <= 1 LSB on >= 99.9 % of pixels and <= 2 LSB everywhere."""
import numpy as np


def depth_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def color_report(gpu_rgba: np.ndarray, ref_rgba: np.ndarray) -> dict:
    d = np.abs(gpu_rgba[..., :3].astype(np.int16) - ref_rgba[..., :3].astype(np.int16)).max(axis=-1)
    n = d.size
    return {"max": int(d.max()) if n else 0, "frac_le1": float((d <= 1).sum() / n) if n else 1.0,
            "n_diff": int((d > 0).sum())}


def assert_color_ok(gpu_rgba, ref_rgba, what=""):
    r = color_report(gpu_rgba, ref_rgba)
    assert r["max"] <= 2 and r["frac_le1"] >= 0.999, (what, r)
    return r


def assert_depth_exact(gpu_depth, ref_depth, what=""):
    if not depth_equal(gpu_depth, ref_depth):
        bad = np.argwhere(gpu_depth.view(np.uint32) != ref_depth.view(np.uint32))
        y, x = bad[0]
        raise AssertionError(f"{what}: {len(bad)} depth mismatches, first at (x={x}, y={y}): "
                             f"gpu={gpu_depth[y, x]!r} ref={ref_depth[y, x]!r}")
