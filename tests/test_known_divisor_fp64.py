"""The FP64 division by a model constant that the adjoint kernels use
(eik_adjoint.hip: cdiv): q = RN(a y), y = RN(1/b), then one fma correction
RN(q + (a - b q) y).  By the FMA division theorems (Markstein) this is RN(a/b)
for every dividend whose quotient neither underflows nor overflows.  Checked
here on the host (glibc's fma is correctly rounded, as the GPU's v_fma_f64
is) bit for bit against IEEE division, for the divisors the kernels use
(dx, dz, 1/dz^2 + 1/dx^2, sqrt(dx^2 + dz^2)) of several spacings, over
random dividends of every magnitude from 2^-900 to 2^900 and both signs,
dividends that are exact multiples of the divisor and their neighbours.
"""
import os
import subprocess

SRC = r"""
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

static double cdiv(double a, double b, double y) {
  const double q = a * y;
  const double r = std::fma(-b, q, a);
  return std::fma(r, y, q);
}

int main() {
  const double sp[][2] = {{10, 10}, {5, 5}, {7, 10}, {10, 3.3}, {1, 0.5}, {12.5, 2.5}, {0.3, 0.7}, {33.3, 1.1}};
  std::mt19937_64 g(12345);
  std::uniform_real_distribution<double> mant(1.0, 2.0);
  std::uniform_int_distribution<int> ex(-900, 900), sign(0, 1);
  long bad = 0, n = 0;
  for (auto& s : sp) {
    const double dx = s[0], dz = s[1];
    const double sum = 1.0 / (dz * dz) + 1.0 / (dx * dx);
    const double divs[4] = {dx, dz, sum, std::sqrt(dx * dx + dz * dz)};
    for (double b : divs) {
      const double y = 1.0 / b;
      for (int k = 0; k < 400000; ++k) {
        double a = std::ldexp(mant(g), ex(g)) * (sign(g) ? -1.0 : 1.0);
        if (k % 4 == 1) a = b * std::ldexp(mant(g), ex(g) / 8);           // near exact multiples
        if (k % 4 == 2) a = std::nextafter(b * std::floor(mant(g) * 1000.0), 0.0);
        const double q1 = cdiv(a, b, y), q2 = a / b;
        ++n;
        if (std::memcmp(&q1, &q2, 8) != 0) ++bad;
      }
      const double z = cdiv(0.0, b, y);
      ++n;
      if (!(z == 0.0 && !std::signbit(z))) ++bad;
    }
  }
  std::printf("%ld %ld\n", n, bad);
  return bad != 0;
}
"""


def test_fma_division_by_constant_is_ieee_division(tmp_path):
    src = os.path.join(tmp_path, "cdiv.cpp")
    exe = os.path.join(tmp_path, "cdiv")
    with open(src, "w") as f:
        f.write(SRC)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe, src], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, text=True, timeout=120)
    n, bad = map(int, r.stdout.split())
    assert n > 10_000_000 and bad == 0, r.stdout
