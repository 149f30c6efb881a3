"""The FP64 elementary functions of the kernels (log for w >= 1, atan2, in
disco_amd/csrc/pair_terms.h), compiled for the host with the same source and
checked against the C library over wide argument ranges: at most a few units
in the last place.  (The device build runs the same code; the FP64 parity tests
check it there through the sums.)"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
// the HIP vector types the header's loopon code names
struct float4 { float x, y, z, w; };
struct double4 { double x, y, z, w; };
#include "pair_terms.h"

struct HostMath {
    static double rcp_d(double x) { return 1.0 / x; }
    static double fabs_d(double x) { return std::fabs(x); }
    static bool signbit_d(double x) { return std::signbit(x); }
    static int exponent_d(double x) { int e; std::frexp(x, &e); return e; }
    static double ldexp_d(double x, int e) { return std::ldexp(x, e); }
};

static double ulps(double got, double want)
{
    if (got == want) return 0.0;
    const double u = std::ldexp(1.0, std::ilogb(want) - 52);
    return std::fabs(got - want) / u;
}

int main()
{
    std::mt19937_64 r(7);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    double worst_log = 0, worst_atan = 0;
    long bad_signs = 0;
    for (int i = 0; i < 2000000; i++) {
        const double w = std::exp(std::fabs(U(r)) * 40.0) * (1.0 + 1e-3 * std::fabs(U(r)));
        worst_log = std::fmax(worst_log, ulps(disco::log_ge1_d<HostMath>(w), std::log(w)));
        const double s = std::ldexp(1.0, (int)(U(r) * 30));
        const double x = U(r) * s, y = U(r) * std::ldexp(s, (int)(U(r) * 8));
        const double a = disco::atan2_own_d<HostMath>(y, x), b = std::atan2(y, x);
        worst_atan = std::fmax(worst_atan, ulps(a, b));
    }
    // quadrants, axes and signed zeros
    const double v[] = {0.0, -0.0, 1.0, -1.0, 0.5, -2.0, 1e-300, -1e300};
    for (double y : v)
        for (double x : v) {
            const double a = disco::atan2_own_d<HostMath>(y, x), b = std::atan2(y, x);
            if (std::signbit(a) != std::signbit(b) || ulps(a, b) > 4) bad_signs++;
        }
    std::printf("%.3f %.3f %ld\n", worst_log, worst_atan, bad_signs);
    return 0;
}
'''


def test_fp64_log_and_atan2_match_the_c_library(tmp_path):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no C++ compiler")
    src = tmp_path / "m.cpp"
    src.write_text(SRC)
    exe = tmp_path / "m"
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "disco_amd", "csrc"),
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    worst_log, worst_atan, bad = float(out[0]), float(out[1]), int(out[2])
    assert worst_log <= 4.0, worst_log
    assert worst_atan <= 4.0, worst_atan
    assert bad == 0
