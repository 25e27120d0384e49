"""The host-side code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/sanitize/san_driver.cpp is built with g++/gcc together with the
library's host sources (the OBJ and TGA parsers, the uniform math, the scene
state) and the oracle, all with -fsanitize=address,undefined
-fno-sanitize-recover=all, and run on valid and malformed OBJ and TGA files
(truncated data, indices past the lists, run-length packets past the image,
absurd dimensions, unknown formats) and on the shipped models where they are
staged.  Any sanitizer report fails the test.  No GPU code is involved."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "hana-softwarerenderer_amd", "csrc")
ASSETS = os.path.join(REPO, "oracle", "_ref", "assets")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                                reason="needs gcc/g++ with the sanitizer runtimes")

FLAGS = ["-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-ffp-contract=off"]


def tga(path, w, h, bpp, rle, data=None, kind=None, truncate=0, bad_run=False):
    """A TGA file in the layout TGAImage reads (18-byte header, then pixels)."""
    if kind is None:
        kind = (11 if rle else 3) if bpp == 1 else (10 if rle else 2)
    hdr = struct.pack("<BBBHHBHHHHBB", 0, 0, kind, 0, 0, 0, 0, 0, w, h, bpp * 8, 0x20)
    if data is None:
        data = np.random.default_rng(w * h + bpp).integers(0, 256, w * h * bpp, dtype=np.uint8).tobytes()
    body = b""
    if rle:
        px = [data[i:i + bpp] for i in range(0, len(data), bpp)]
        i = 0
        while i < len(px):
            n = min(128, len(px) - i)
            body += bytes([n - 1]) + b"".join(px[i:i + n])
            i += n
        if bad_run:
            body = bytes([0x80 | 127]) + px[0]   # a run of 128 pixels past a smaller image
    else:
        body = data
    if truncate:
        body = body[:-truncate]
    with open(path, "wb") as f:
        f.write(hdr + body)


def obj(path, text):
    with open(path, "w") as f:
        f.write(text)


SPHERE_OBJ = None


def sphere_obj(n=12):
    lines = []
    for i in range(n + 1):
        th = np.pi * i / n
        for j in range(n):
            ph = 2 * np.pi * j / n
            p = (np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph))
            lines.append("v %.6f %.6f %.6f" % p)
            lines.append("vt %.6f %.6f" % (j / n, i / n))
            lines.append("vn %.6f %.6f %.6f" % p)
    for i in range(n):
        for j in range(n):
            a, b = i * n + j + 1, i * n + (j + 1) % n + 1
            c, d = a + n, b + n
            lines.append(f"f {a}/{a}/{a} {c}/{c}/{c} {b}/{b}/{b}")
            lines.append(f"f {b}/{b}/{b} {c}/{c}/{c} {d}/{d}/{d}")
    return "\n".join(lines) + "\n"


def test_host_code_under_sanitizers(tmp_path):
    exe = tmp_path / "san_driver"
    objs = []
    for src in ("hana_assets.cpp", "hana_hostmath.cpp", "hana_scene.cpp"):
        o = tmp_path / (src + ".o")
        subprocess.run(["g++", "-std=c++17", *FLAGS, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c",
                        os.path.join(CSRC, src), "-o", str(o)], check=True)
        objs.append(str(o))
    o = tmp_path / "hana_oracle.o"
    subprocess.run(["gcc", "-std=c11", *FLAGS, "-c", os.path.join(REPO, "oracle", "hana_oracle.c"), "-o", str(o)],
                   check=True)
    objs.append(str(o))
    subprocess.run(["g++", "-std=c++17", *FLAGS, "-I" + os.path.join(REPO, "include"), "-I" + os.path.join(REPO, "oracle"),
                    os.path.join(REPO, "tests", "sanitize", "san_driver.cpp"), *objs, "-lm", "-o", str(exe)], check=True)
    files = []
    d = tmp_path / "in"
    d.mkdir()
    # OBJ: valid, then malformed
    obj(d / "sphere.obj", sphere_obj())
    obj(d / "quad.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvn 0 0 1\n"
                        "f 1/1/1 2/2/1 3/3/1 4/1/1\n# comment\n\n")
    obj(d / "past_list.obj", "v 0 0 0\nv 1 0 0\nvt 0 0\nvn 0 0 1\nf 1/1/1 2/1/1 7/1/1\n")
    obj(d / "negative.obj", "v 0 0 0\nvt 0 0\nvn 0 0 1\nf -1/1/1 1/1/1 1/1/1\n")
    obj(d / "short_face.obj", "v 0 0 0\nv 1 0 0\nvt 0 0\nvn 0 0 1\nf 1/1/1 2/1/1\n")
    obj(d / "no_slashes.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    obj(d / "garbage.obj", "v 1e999 nan x\nvt\nvn 1 2\nf a/b/c d/e/f g/h/i\n" + "v " * 1000 + "\n")
    obj(d / "empty.obj", "")
    # TGA: valid in every format, then malformed
    for bpp in (1, 3, 4):
        for rle in (False, True):
            tga(d / f"ok_{bpp}_{int(rle)}.tga", 37, 23, bpp, rle)
    tga(d / "truncated_raw.tga", 40, 40, 3, False, truncate=100)
    tga(d / "truncated_rle.tga", 40, 40, 4, True, truncate=7)
    tga(d / "run_past_image.tga", 4, 4, 3, True, bad_run=True)
    tga(d / "bad_bpp.tga", 8, 8, 2, False, data=bytes(8 * 8 * 2))
    tga(d / "bad_kind.tga", 8, 8, 3, False, kind=9)
    tga(d / "huge.tga", 65535, 65535, 4, False, data=b"\x00" * 64)
    (d / "tiny.tga").write_bytes(b"\x00" * 7)
    files = sorted(str(p) for p in d.iterdir())
    for m in ("african_head", "diablo3_pose"):
        base = os.path.join(ASSETS, m, m)
        if os.path.exists(base + ".obj"):
            files += [base + ".obj", base + "_diffuse.tga", base + "_nm_tangent.tga"]
    env = dict(os.environ, SAN_TMP=str(tmp_path), ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe), *files], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    line = r.stdout.strip().splitlines()[-1]
    assert "0 failures" in line, line
    loaded = int(line.split()[1])
    assert loaded >= 2   # sphere and quad at least
