"""The elementary functions of the kernels (FP64 log for w >= 1, atan2 and the
walks' 2 atanh, FP32 atan2, in disco_amd/csrc/pair_terms.h), compiled for the host with the
same source and checked against the C library over wide argument ranges: at
most a few units in the last place.  (The device build runs the same code, with
the hardware reciprocal in place of the host's division; the parity tests check
it there through the sums.)"""
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
    static float rcp_f(float x) { return 1.0f / x; }
    static float fabs_f(float x) { return std::fabs(x); }
    static float fmax_f(float a, float b) { return std::fmax(a, b); }
    static float fmin_f(float a, float b) { return std::fmin(a, b); }
    static float copysign_f(float a, float b) { return std::copysign(a, b); }
    static double rcp_d(double x) { return 1.0 / x; }
    static double fabs_d(double x) { return std::fabs(x); }
    static bool signbit_d(double x) { return std::signbit(x); }
    static int exponent_d(double x) { int e; std::frexp(x, &e); return e; }
    static double ldexp_d(double x, int e) { return std::ldexp(x, e); }
    static bool any(bool p) { return p; }      // one lane: each z takes its own branch
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
    // FP32 atan2 against the double library value rounded to float's ulp, over
    // the magnitudes the walks hand it (1e-18 .. 1e36) and every angle
    double worst_f = 0;
    std::uniform_real_distribution<double> A(-3.14159265358979, 3.14159265358979);
    for (int i = 0; i < 4000000; i++) {
        const double m = std::pow(10.0, -18.0 + 54.0 * (0.5 + 0.5 * U(r)));
        const double an = (i & 1) ? A(r) : std::ldexp(U(r), -(int)(24.0 * std::fabs(U(r))));
        const float x = (float)(m * std::cos(an)), y = (float)(m * std::sin(an));
        const double want = std::atan2((double)y, (double)x);
        if (want == 0.0) continue;
        const double u = std::ldexp(1.0, std::ilogb(want) - 23);
        worst_f = std::fmax(worst_f, std::fabs(disco::atan2_own_f<HostMath>(y, x) - want) / u);
    }
    for (float y : {0.0f, -0.0f, 1.0f, -1.0f, 3e-18f})
        for (float x : {0.0f, 1.0f, -1.0f, 2e30f, -5e-10f}) {
            const float a = disco::atan2_own_f<HostMath>(y, x);
            const double b = std::atan2((double)y, (double)x);
            if (x == 0.0f && y == 0.0f) { if (a != 0.0f || std::signbit(a) != std::signbit(y)) bad_signs++; continue; }
            if (std::signbit(a) != std::signbit(b) || std::fabs(a - b) > 4e-7 * std::fmax(std::fabs(b), 1e-30)) bad_signs++;
        }
    // 2 atanh(z) of the FP64 walks over its three branches (the series to z^13
    // below 1/16, to z^27 below 1/4, the log form above), against the C library
    // in long double
    double worst_at = 0;
    for (int i = 0; i < 2000000; i++) {
        const double z = (i & 1) ? 0.0625 * (0.5 + 0.5 * U(r))
                                 : std::ldexp(0.5 + 0.5 * U(r), -(int)(40.0 * std::fabs(U(r))));
        const double zz = (i % 3 == 0) ? 0.999 * (0.5 + 0.5 * U(r)) : z;
        const long double want = 2.0L * std::atanh((long double)zz);
        const double got = disco::two_atanh_vote_d<HostMath>(zz);
        worst_at = std::fmax(worst_at, (double)(std::fabs((long double)got - want) /
                                                 std::ldexp(1.0L, std::ilogb((double)want) - 52)));
    }
    std::printf("%.3f %.3f %ld %.3f %.3f\n", worst_log, worst_atan, bad_signs, worst_f, worst_at);
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
    worst_f, worst_at = float(out[3]), float(out[4])
    assert worst_at <= 4.0, worst_at
    assert worst_log <= 4.0, worst_log
    assert worst_atan <= 4.0, worst_atan
    assert worst_f <= 3.0, worst_f
    assert bad == 0
