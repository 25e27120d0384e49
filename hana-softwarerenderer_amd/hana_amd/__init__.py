"""hana_amd: MI355X (gfx950) rasterisation pipeline of Hana-SoftwareRenderer.

The compute path is libhana_gpu.so (hand-written HIP kernels behind the C ABI
of include/hana_gpu.h); this package holds its ctypes binding (gpu), seeded
workload generators (scenegen) and the serving loop (serve).
"""
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPO_DIR = os.path.dirname(PKG_DIR)
