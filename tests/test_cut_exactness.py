"""The kernels decide the reference's per-axis kernel cut-off, abs(dx/h) > 3
with a rounded division (sph_lib.cc:217-220), by the division-free identity
fma(-3, h, |dx|) > 2^-52 h.  Check it against real IEEE division on the
values that matter: |dx| within a few ulp of 3h (both sides, ties included)
for h across many binades, plus random pairs.  Compiled with gcc on the CPU.
"""
import os
import subprocess
import tempfile

SRC = r"""
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
static uint64_t st = 88172645463325252ull;
static double urand(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (st >> 11) * (1.0 / 9007199254740992.0); }
static int fma_out(double a, double h) { return fma(-3.0, h, a) > ldexp(1.0, -52) * h; }
static int div_out(double a, double h) { return fabs(a / h) > 3.0; }
int main(void) {
  long checked = 0, bad = 0, outs = 0;
  for (int it = 0; it < 200000; it++) {
    double h = ldexp(1.0 + urand(), (int)(urand() * 80) - 40);
    double b = 3.0 * h;
    double a = b;
    for (int k = 0; k < 6; k++) a = nextafter(a, 0.0);
    for (int k = 0; k < 13; k++) {
      int f = fma_out(a, h), d = div_out(a, h);
      checked++; outs += d;
      if (f != d) { bad++; if (bad < 5) printf("mismatch a=%a h=%a fma=%d div=%d\n", a, h, f, d); }
      a = nextafter(a, INFINITY);
    }
    /* exact midpoint between 3 and 3 + 2^-51 in units of h, when representable */
    double m = h * (3.0 + ldexp(1.0, -52));
    if (fabs(m / h - 3.0) <= ldexp(1.0, -51)) {
      checked++;
      if (fma_out(m, h) != div_out(m, h)) bad++;
    }
    double r = urand() * 6.0 * h;
    checked++;
    if (fma_out(r, h) != div_out(r, h)) bad++;
  }
  printf("checked %ld bad %ld outside %ld\n", checked, bad, outs);
  return bad != 0;
}
"""


def test_fma_cut_matches_division():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "cut.c")
        exe = os.path.join(d, "cut")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
        print(r.stdout)
        assert r.returncode == 0, r.stdout
        # both outcomes must actually occur near the boundary
        outs = int(r.stdout.split("outside")[1])
        assert outs > 100000

