"""glibc's pow(x, 2.0) is the correctly rounded square x * x for every x the
bond yield function can meet (ForceLaw_StickedLinks.cpp:143-144 evaluates
pow(|ft| / ft0, power) with power 2 in every deck here).  The device computes
that case as x * x, so its break decisions are the reference's bit for bit.
Checked on 10^8 doubles whose exponents span 2^-60 .. 2^60 (every mantissa
pattern equally likely), and on the squares' rounding midpoints' neighbours,
where a pow with error above half an ulp would round the other way."""
import os
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 88172645463325252ull;
static uint64_t xr(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static int check(double x) { volatile double p = pow(x, 2.0), q = x * x; return p != q; }
int main(void) {
  long bad = 0, n = 100000000;
  for (long k = 0; k < n; k++) {
    uint64_t e = (uint64_t)(1023 - 60 + (xr() >> 57) % 121) << 52;
    uint64_t bits = e | (xr() & 0xfffffffffffffull);
    double x;
    memcpy(&x, &bits, 8);
    bad += check(x);
  }
  /* x = 1 + k 2^-26 (k odd): x^2 = 1 + k 2^-25 + k^2 2^-52 has its 53rd-bit
     tail exactly at a rounding midpoint when k^2 is odd at bit 2^-53 */
  for (long k = 1; k < 20000000; k += 2) {
    double x = 1.0 + (double)k * 0x1p-26;
    bad += check(x) + check(nextafter(x, 2.0)) + check(nextafter(x, 0.0));
  }
  printf("%ld\n", bad);
  return 0;
}
"""


def test_glibc_pow_square_is_x_times_x(tmp_path):
    src = tmp_path / "powsq.c"
    exe = tmp_path / "powsq"
    src.write_text(SRC)
    try:
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True,
                       capture_output=True)
    except (OSError, subprocess.CalledProcessError) as exc:
        pytest.skip("no C compiler: %s" % exc)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=300).stdout
    assert int(out.strip()) == 0
