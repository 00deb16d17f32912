# SPDX-License-Identifier: Apache-2.0
"""The device's atan2f (nph_math.h: nph_atan2f, fdlibm's float method) against
the C library's, which the reference calls (scalar.h:239-242): every atanf
input on a stride through all 2^32 bit patterns, a seeded sample of
atan2f pairs (random bits and ordinary ratios), and the special operands.
Compiled with gcc from the same header the kernels include."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include "nph_math.h"
#include <math.h>
#include <stdio.h>
int main(void) {
    unsigned long long bad = 0, n = 0;
    for (unsigned long long u = 0; u < (1ULL << 32); u += 61) {
        const float y = nph_u2f((unsigned)u);
        const float a = atan2f(y, 1.0f), b = nph_atan2f(y, 1.0f);
        n++;
        if (nph_f2u(a) != nph_f2u(b) && !(isnan(a) && isnan(b))) bad++;
    }
    for (unsigned long long t = 0; t < 20000000ULL; ++t) {
        unsigned long long s = t * 6364136223846793005ULL + 1442695040888963407ULL;
        s ^= s >> 29; s *= 0xbf58476d1ce4e5b9ULL; s ^= s >> 32;
        float y = nph_u2f((unsigned)s), x = nph_u2f((unsigned)(s >> 32));
        if (t & 1) {
            y = (float)((int)(s & 0xffff) - 32768) / 1000.0f;
            x = (float)((int)((s >> 16) & 0xffff) - 32768) / 997.0f;
        }
        const float a = atan2f(y, x), b = nph_atan2f(y, x);
        n++;
        if (nph_f2u(a) != nph_f2u(b) && !(isnan(a) && isnan(b))) bad++;
    }
    const float sp[] = {0.0f, -0.0f, 1.0f, -1.0f, INFINITY, -INFINITY, 1e-30f, -1e-30f, 1e30f,
                        -1e30f, 3.0e-39f, -3.0e-39f};
    for (unsigned i = 0; i < 12; ++i)
        for (unsigned j = 0; j < 12; ++j) {
            const float a = atan2f(sp[i], sp[j]), b = nph_atan2f(sp[i], sp[j]);
            n++;
            if (nph_f2u(a) != nph_f2u(b)) bad++;
        }
    printf("%llu %llu\n", bad, n);
    return 0;
}
"""


def test_atan2f_matches_the_c_library(tmp_path):
    src = tmp_path / "atan2.c"
    src.write_text(HARNESS)
    exe = tmp_path / "atan2"
    subprocess.run(["gcc", "-std=c17", "-O2", "-ffp-contract=off", "-msse4.2",
                    "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "nepheline_amd", "csrc"), "-o", str(exe), str(src),
                    "-lm"], check=True, capture_output=True)
    bad, n = map(int, subprocess.run([str(exe)], check=True, capture_output=True,
                                     text=True).stdout.split())
    assert n > 90_000_000
    assert bad == 0
