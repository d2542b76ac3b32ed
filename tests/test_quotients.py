"""The quotient identity the kernels use in place of IEEE division
(dasha_amd/csrc/mhd_device.h: divc, div3, div6): with y = RN(1/b),

    q0 = RN(a y),  q = RN(q0 + RN-exact(a - q0 b) y)   (two FMAs)

is the correctly rounded a / b for finite a, b without overflow or underflow
(Markstein's single-correction theorem: y within half an ulp of 1/b, q0
within one ulp of a/b).  The kernels divide by rho through one reciprocal
(flux tensor, v = m/rho, Hall term, setdt) and by 3 and 6 in the RK4 register
updates (t3dsdsc.f:1090-1132), so every such quotient must equal the
reference's division bit for bit.  The restatement below is compiled with gcc
(-ffp-contract=off, fma from libm) and compared with the C division on random
operands over a wide exponent range; dropping the correction (q = q0) is shown
to differ, so the check can fail."""
import os
import shutil
import subprocess

import pytest

_SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t nx(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double rnd(int span) {
    uint64_t u = nx();
    uint64_t e = (uint64_t)(1023 - span + (int)((u >> 52) % (uint64_t)(2 * span)));
    uint64_t b = (u & 0x800FFFFFFFFFFFFFull) | (e << 52);
    double d; memcpy(&d, &b, 8); return d;
}
static double divc(double a, double b, double y) { double q0 = a * y; return fma(fma(-q0, b, a), y, q0); }
static double div3(double a) { return divc(a, 3.0, 0x1.5555555555555p-2); }
static double div6(double a) { return divc(a, 6.0, 0x1.5555555555555p-3); }
int main(int argc, char **argv) {
    long n = atol(argv[1]), bad_gen = 0, bad36 = 0, plain = 0;
    if (0x1.5555555555555p-2 != 1.0 / 3.0 || 0x1.5555555555555p-3 != 1.0 / 6.0) { puts("constants"); return 2; }
    for (long i = 0; i < n; i++) {
        double a = rnd(400), b = rnd(400), y = 1.0 / b;
        if (divc(a, b, y) != a / b) bad_gen++;
        if (a * y != a / b) plain++;
        double c = rnd(600);
        if (div3(c) != c / 3.0 || div6(c) != c / 6.0) bad36++;
    }
    printf("%ld %ld %ld\n", bad_gen, bad36, plain);
    return 0;
}
"""


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    d = tmp_path_factory.mktemp("quot")
    src = d / "q.c"
    src.write_text("#include <stdlib.h>\n" + _SRC)
    exe = d / "q"
    subprocess.run([cc, "-O2", "-ffp-contract=off", str(src), "-o", str(exe), "-lm"], check=True)
    return str(exe)


def test_reciprocal_quotient_is_the_ieee_quotient(prog):
    out = subprocess.run([prog, "20000000"], check=True, capture_output=True, text=True).stdout.split()
    bad_gen, bad36, plain = (int(v) for v in out)
    assert bad_gen == 0, "divc differs from a / b"
    assert bad36 == 0, "div3 / div6 differ from a / 3, a / 6"
    # the uncorrected product a * RN(1/b) is wrong often enough to be seen
    assert plain > 1000
