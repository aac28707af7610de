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


SRC_F32 = r"""
#include <math.h>
#include <stdio.h>
#include <stdint.h>
static uint64_t st = 0x9E3779B97F4A7C15ull;
static double urand(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (st >> 11) * (1.0 / 9007199254740992.0); }
/* the FP64 pair predicate of the kernels: kernel box at h_ij, list rule, support */
static int exact_in(double xi[3], double xj[3], double hi, double hj, double lf, double supp, int same) {
  double a = 0;
  for (int k = 0; k < 3; k++) a = fmax(a, fabs(xi[k] - xj[k]));
  double h = 0.5 * (hi + hj);
  if (fma(-3.0, h, a) > ldexp(1.0, -52) * h) return 0;
  return (same || a <= lf) && a <= supp;
}
int main(void) {
  long checked = 0, bad = 0, amb = 0, near = 0;
  for (int it = 0; it < 400000; it++) {
    double hmax = ldexp(1.0 + urand(), (int)(urand() * 24) - 6);
    double E = 3.0 * hmax * (1 + 1e-9);
    int r = 1 + (int)(urand() * 2), span = 1 + (int)(urand() * 4);
    int nx = span + 2 * r, ny = 2 * r + 1;
    double dlim = (double)((nx > ny ? nx : ny) + 1) * E;
    double o[3], xi[3], xj[3];
    for (int k = 0; k < 3; k++) o[k] = (urand() - 0.5) * ldexp(1.0, 20) * hmax;
    double hi = hmax * (0.5 + 0.5 * urand()), hj = hmax * (0.5 + 0.5 * urand());
    double lf = 4.25 * hi * (0.8 + 0.4 * urand()), supp = 8.0 * hi;
    int same = urand() < 0.3;
    for (int k = 0; k < 3; k++) xi[k] = o[k] + (urand() - 0.5) * span * E;
    /* put the neighbour on one of the decision surfaces, or anywhere */
    double thr[3] = {1.5 * (hi + hj), lf, supp};
    double target = thr[(int)(urand() * 3)] * (1 + (urand() - 0.5) * ldexp(1.0, -(int)(urand() * 40)));
    int ax = (int)(urand() * 3);
    for (int k = 0; k < 3; k++) xj[k] = xi[k] + (urand() - 0.5) * 2 * target;
    xj[ax] = xi[ax] + (urand() < 0.5 ? -target : target);
    if (urand() < 0.2) for (int k = 0; k < 3; k++) xj[k] = o[k] + (urand() - 0.5) * 2 * dlim;
    int far = 0;
    for (int k = 0; k < 3; k++) far |= !(fabs(xj[k] - o[k]) < dlim) || !(fabs(xi[k] - o[k]) < dlim);
    if (far) continue;
    int ex = exact_in(xi, xj, hi, hj, lf, supp, same);
    /* FP32 decision of the kernels (with and without fused multiply-add) */
    float xif[3], xjf[3];
    for (int k = 0; k < 3; k++) { xif[k] = (float)(xi[k] - o[k]); xjf[k] = (float)(xj[k] - o[k]); }
    float am = fmaxf(fabsf(xif[0] - xjf[0]), fmaxf(fabsf(xif[1] - xjf[1]), fabsf(xif[2] - xjf[2])));
    float m = (float)(dlim * ldexp(1.0, -19));
    float hk = 0.5f * ((float)hi + (float)hj);
    float lff = (float)lf, sup = (float)supp;
    for (int v = 0; v < 2; v++) {
      float tlo = v ? fmaf(3.0f, hk, -m) : 3.0f * hk - m;
      float thi = v ? fmaf(3.0f, hk, m) : 3.0f * hk + m;
      int in = am <= tlo && (same || am <= lff - m) && am <= sup - m;
      int out = am > thi || (!same && am > lff + m) || am > sup + m;
      checked++;
      if (in && out) bad++;
      if (in && !ex) { bad++; if (bad < 5) printf("false in: am=%a ex=%d\n", am, ex); }
      if (out && ex) { bad++; if (bad < 5) printf("false out: am=%a ex=%d\n", am, ex); }
      if (!in && !out) amb++;
    }
    near++;
  }
  printf("checked %ld bad %ld ambiguous %ld cases %ld\n", checked, bad, amb, near);
  return bad != 0;
}
"""


def test_fp32_prefilter_never_contradicts_exact_predicate():
    """The pair kernels decide most candidates in FP32 on chunk-relative
    coordinates and re-run the FP64 test only inside the rounding margin;
    a sure FP32 decision must agree with the FP64 predicate."""
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "f32.c")
        exe = os.path.join(d, "f32")
        with open(src, "w") as f:
            f.write(SRC_F32)
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
        print(r.stdout)
        assert r.returncode == 0, r.stdout
        amb = int(r.stdout.split("ambiguous")[1].split()[0])
        checked = int(r.stdout.split("checked")[1].split()[0])
        assert checked > 400000 and 0 < amb < checked
