"""The raster's coverage quotients n / u.z from one refined reciprocal.

k_raster computes the three barycentric quotients of a candidate pixel as
q = n r; e = fma(-u.z, q, n); q = fma(e, r, q); e = fma(-u.z, q, n);
q = fma(e, r, q), with r the reciprocal of u.z from the hardware's rcp
(within 1 ulp) refined once, r = fma(fma(-u.z, r0, 1), r0, r0).  Those are
the operations of the hardware's IEEE division sequence when its scaling
steps do nothing, which the kernel ensures by taking the sequence only for
2^-7 < |u.z| <= 2^45 and |n| >= 2^-80 (or n == 0).  On the GPU the images are
bit-identical either way; this pins the arithmetic itself on the host: for
every reciprocal within 1 ulp (the correctly rounded one and its two
neighbours) the sequence returns the correctly rounded quotient, over random
and structured operands of that range.  The shading's byte / 255 (texels,
the R8 shadow map) is one product with RN(1/255) and one fma correction;
that equals the division for every byte.  Host C (gcc), fmaf from libm, no
GPU."""
import os
import shutil
import subprocess

import pytest

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint64_t nxt(void) {   /* splitmix64 */
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static float bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t ubits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
/* a float of random mantissa and sign with exponent in [lo, hi] */
static float rnd(int lo, int hi) {
    uint64_t r = nxt();
    int e = lo + (int)(r % (uint64_t)(hi - lo + 1));
    uint32_t m = (uint32_t)(r >> 20) & 0x7FFFFFu;
    uint32_t s = (uint32_t)(r >> 63) << 31;
    return bits(s | ((uint32_t)(e + 127) << 23) | m);
}
static volatile float sink;
static float quot(float n, float uz, float r0) {
    float e0 = fmaf(-uz, r0, 1.0f);
    float r = fmaf(e0, r0, r0);
    float q = n * r;
    float e = fmaf(-uz, q, n);
    q = fmaf(e, r, q);
    e = fmaf(-uz, q, n);
    return fmaf(e, r, q);
}
static long bad = 0, checked = 0;
static void check(float n, float uz) {
    float exact = n / uz;
    float r = 1.0f / uz;
    float rs[3] = {r, nextafterf(r, INFINITY), nextafterf(r, -INFINITY)};
    for (int k = 0; k < 3; k++) {
        float q = quot(n, uz, rs[k]);
        checked++;
        /* zero quotients may differ in sign only */
        if (n == 0.0f ? q != 0.0f : ubits(q) != ubits(exact)) {
            if (bad < 5) printf("MISMATCH n=%a uz=%a r=%a q=%a exact=%a\n", n, uz, rs[k], q, exact);
            bad++;
        }
    }
}
int main(int argc, char** argv) {
    long n_random = argc > 1 ? atol(argv[1]) : 2000000;
    for (long i = 0; i < n_random; i++) {
        float uz = rnd(-6, 44);                 /* 2^-6 <= |u.z| < 2^45 (0.01 < 2^-6 would pass too) */
        if (fabsf(uz) <= 0.01f) continue;
        check(rnd(-80, 46), uz);                /* 2^-80 <= |n| < 2^47 */
    }
    /* structured: quotients at and next to exact values, near 1, mantissa extremes */
    for (long i = 0; i < n_random / 4; i++) {
        float uz = rnd(-6, 40);
        float k = (float)(int32_t)(nxt() % 2000001) - 1000000.0f;
        float n = uz * k;                       /* exact multiples where representable */
        check(n, uz);
        check(nextafterf(n, INFINITY), uz);
        check(nextafterf(n, -INFINITY), uz);
        check(uz, uz);
        check(nextafterf(uz, 0.0f), uz);
        uint32_t ones = (ubits(uz) & 0xFF800000u) | 0x7FFFFFu;   /* all-ones mantissa divisor */
        check(rnd(-80, 46), bits(ones));
        check(0.0f, uz);
        check(-0.0f, uz);
    }
    printf("checked %ld bad %ld\n", checked, bad);
    return bad != 0;
}
"""


def test_refined_reciprocal_quotients_are_correctly_rounded(tmp_path):
    c = tmp_path / "cover_div.c"
    c.write_text(SRC.replace("#include <string.h>", "#include <string.h>\n#include <stdlib.h>"))
    exe = tmp_path / "cover_div"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-o", str(exe), str(c), "-lm"], check=True)
    out = subprocess.run([str(exe), "3000000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:]
    assert "bad 0" in out.stdout


BYTE_SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
int main(void) {
    const float r = 0x1.010102p-8f;
    int bad = 0;
    if (r != 1.0f / 255.0f) { printf("constant is not RN(1/255)\n"); return 1; }
    for (int b = 0; b < 256; b++) {
        float x = (float)b, q = x * r, e = fmaf(-255.0f, q, x), d = fmaf(e, r, q), ex = x / 255.0f;
        uint32_t u, v;
        memcpy(&u, &d, 4);
        memcpy(&v, &ex, 4);
        if (u != v) { printf("byte %d: %a against %a\n", b, d, ex); bad++; }
    }
    printf("bad %d\n", bad);
    return bad != 0;
}
"""


def test_byte_over_255_is_the_division(tmp_path):
    c = tmp_path / "byte_div.c"
    c.write_text(BYTE_SRC)
    exe = tmp_path / "byte_div"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-o", str(exe), str(c), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "bad 0" in out.stdout, out.stdout


LEMMA_SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t st = 0xD1B54A32D192ED03ull;
static uint64_t nxt(void) {   /* splitmix64 */
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static float bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t ubits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
/* an input the per-triangle test admits: zero, or 2^-8 <= |v| <= 2^22, drawn
   near the ends of that range, near integers and near 1 - 2^-24 */
static float input(void) {
    uint64_t r = nxt();
    uint32_t s = (uint32_t)(r >> 63) << 31, m = (uint32_t)(r >> 20) & 0x7FFFFFu;
    switch (r % 7) {
    case 0: return 0.0f;
    case 1: return bits(s | ((uint32_t)(127 - 8) << 23) | m);              /* [2^-8, 2^-7) */
    case 2: return bits(s | ((uint32_t)(127 - 8) << 23) | (m & 0xF));      /* just above 2^-8 */
    case 3: return bits(s | ((uint32_t)(127 + 21) << 23) | m);             /* [2^21, 2^22) */
    case 4: { float k = (float)(int32_t)(nxt() % 1000); int d = (int)(nxt() % 9) - 4;
              float v = k; for (int i = 0; i < abs(d); i++) v = nextafterf(v, d < 0 ? -INFINITY : INFINITY);
              return fabsf(v) >= 0x1p-8f ? v : 0.0f; }                    /* integers and their neighbours */
    case 5: return (r & 1) ? 1.0f - 0x1p-24f : -(1.0f - 0x1p-24f);
    default: { int e = (int)(nxt() % 30) - 8; return bits(s | ((uint32_t)(127 + e) << 23) | m); }
    }
}
int main(int argc, char** argv) {
    long n = argc > 1 ? atol(argv[1]) : 1000000, bad = 0, nonzero = 0;
    float least = INFINITY;
    for (long i = 0; i < n; i++) {
        float ax = input(), ay = input(), s0x = input(), s0y = input(), s1x = input(), s1y = input();
        if (fabsf(ax) > 0x1p22f || fabsf(ay) > 0x1p22f) continue;
        for (int k = 0; k < 8; k++) {
            /* pixels at and next to the vertex, and anywhere */
            long px = k < 4 ? lroundf(ax) + (long)(nxt() % 3) - 1 : (long)(nxt() % (1u << 20));
            long py = k < 4 ? lroundf(ay) + (long)(nxt() % 3) - 1 : (long)(nxt() % (1u << 20));
            if (px < 0 || py < 0 || px > (1 << 20) || py > (1 << 20)) continue;
            float Px = (float)px, Py = (float)py;
            float s0z = ax - Px, s1z = ay - Py;
            float p1 = s0y * s1z, p2 = s0z * s1y, p3 = s0z * s1x, p4 = s0x * s1z;
            float ux = p1 - p2, uy = p3 - p4, sum = ux + uy;
            float v[3] = {ux, uy, sum};
            for (int j = 0; j < 3; j++) {
                if (v[j] == 0.0f) continue;
                nonzero++;
                if (fabsf(v[j]) < least) least = fabsf(v[j]);
                if (fabsf(v[j]) < 0x1p-62f) {
                    if (bad < 5) printf("SMALL %a from %a %a %a %a %a %a at %ld %ld\n", v[j], ax, ay, s0x, s0y, s1x, s1y, px, py);
                    bad++;
                }
            }
        }
    }
    printf("nonzero %ld least %a bad %ld\n", nonzero, least, bad);
    return bad != 0;
}
"""


def test_numerators_of_admitted_triangles_are_not_tiny(tmp_path):
    """The raster skips the per-candidate range check for a triangle whose
    inputs (a's screen position and both edges) are each zero or at least 2^-8
    in magnitude: then every nonzero numerator u.x, u.y, u.x + u.y at an integer
    pixel is at least 2^-62, inside the range where the shared-reciprocal
    quotient equals the division (the test above).  Adversarial inputs: the
    ends of the admitted range, integers and their neighbours, 1 - 2^-24, and
    pixels at and next to the vertex."""
    c = tmp_path / "lemma.c"
    c.write_text(LEMMA_SRC)
    exe = tmp_path / "lemma"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-o", str(exe), str(c), "-lm"], check=True)
    out = subprocess.run([str(exe), "2000000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and "bad 0" in out.stdout, out.stdout[-2000:]
