"""The C ABI library: builds, loads, exports every declared entry point, and
its host-side uniform math equals the reference's bit for bit (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from hana_amd import scenegen as sg
from hana_amd.gpu import LIB_PATH, Uniforms, build_uniforms, load_library
from pyoracle import CPURenderer

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "hana_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(hana_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "libhana_gpu.so is not built"
    lib = C.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    load_library()   # every ctypes declaration resolves


def test_uniform_struct_matches_header():
    # 6 matrices, view_pos, light_dir, 4 colours, gloss, bump_scale, flag, pad
    assert C.sizeof(Uniforms) == 6 * 64 + 2 * 12 + 4 * 16 + 4 * 4


@pytest.mark.parametrize("seed", range(12))
def test_host_uniforms_bit_exact(seed):
    """hana_build_uniforms == the oracle's DrawModel::draw uniforms (and the
    compiled reference's, where built) for seeded cameras and lights."""
    rng = np.random.default_rng(seed)
    sc = sg.small_scene(sg.NORMALMAP, shadow=bool(seed % 2), seed=seed,
                        width=int(rng.integers(1, 2000)), height=int(rng.integers(1, 2000)),
                        cam=tuple(float(np.float32(x)) for x in rng.uniform(-3, 3, 3)))
    sc.obj_position = tuple(float(np.float32(x)) for x in rng.uniform(-1, 1, 3))
    sc.obj_rotation = tuple(float(np.float32(x)) for x in rng.uniform(-180, 180, 3))
    sc.obj_scale = tuple(float(np.float32(x)) for x in rng.uniform(0.3, 2, 3))
    u = build_uniforms(sc)
    assert bytes(u) == bytes(CPURenderer("oracle").build_uniforms(sc))
    from pyoracle import REF_SO
    if os.path.exists(REF_SO):
        assert bytes(u) == bytes(CPURenderer("reference").build_uniforms(sc))


def test_invalid_arguments_are_rejected():
    lib = load_library()
    f = np.zeros(3, np.float32)
    fp = f.ctypes.data_as(C.POINTER(C.c_float))
    u = Uniforms()
    assert lib.hana_build_uniforms(fp, fp, 1.0, fp, fp, fp, fp, 0, 10, 0, C.byref(u)) == -1
    assert lib.hana_build_uniforms(None, fp, 1.0, fp, fp, fp, fp, 10, 10, 0, C.byref(u)) == -1
    assert lib.hana_gpu_create(0, None) == -1
    assert lib.hana_gpu_destroy(None) == -1
    assert lib.hana_draw(None, None) == -1
    assert lib.hana_render_frame(None, None) == -1
    assert lib.hana_gpu_last_error(None) == b"no context"


def test_header_is_c99_and_links(tmp_path):
    """include/hana_gpu.h compiles as C99 and C++17, and a C program using the
    documented calls links against libhana_gpu.so (run only as far as the
    host-side entry points, which need no device)."""
    import os
    import subprocess
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    inc = os.path.join(repo, "include")
    pkg = os.path.join(repo, "hana-softwarerenderer_amd")
    src = tmp_path / "use.c"
    src.write_text(r"""
#include <stdio.h>
#include <string.h>
#include "hana_gpu.h"
int main(void) {
    float cam[3] = {0, 0, 2}, tgt[3] = {0, 0, 0}, light[3] = {2, 2, 2};
    float pos[3] = {0, 0, 0}, rot[3] = {0, 0, 0}, scl[3] = {1, 1, 1};
    hana_uniforms u;
    if (hana_build_uniforms(cam, tgt, 1.5f, light, pos, rot, scl, 96, 64, 1, &u) != HANA_OK) return 1;
    hana_image img;
    if (hana_tga_load("/nonexistent.tga", 0, &img) != HANA_E_INVALID) return 2;
    if (strstr(hana_asset_last_error(), "cannot open") == NULL) return 3;
    if (hana_draw(NULL, NULL) != HANA_E_INVALID) return 4;
    printf("%d %.6f\n", (int)sizeof(hana_uniforms), (double)u.camera_vp[0]);
    return 0;
}
""")
    exe = tmp_path / "use"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", f"-I{inc}", str(src), "-o", str(exe),
                    f"-L{pkg}", "-lhana_gpu", f"-Wl,-rpath,{pkg}"], check=True)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-x", "c++", f"-I{inc}",
                    str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert int(out[0]) == 6 * 64 + 2 * 12 + 4 * 16 + 4 * 4
