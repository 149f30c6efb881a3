"""Host side of the line-plan packer (disco_amd/csrc/plan_pack.h), compiled for
the host from the same source: which loops take the frame walks (simple flat
polygons: their solid angle is the principal argument of the fan product) and
which the general walk with the winding count (off the plane, self-crossing,
winding twice, lobes of opposite sense), the normal of a contour with no vector
area, and the per-segment constants of the frame form (edge, 1/L, w3 t)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include <cmath>
#include <cstdio>
#include <vector>
#define DISCO_HD inline
struct float4 { float x, y, z, w; };
struct double4 { double x, y, z, w; };
#include "plan_pack.h"
using namespace disco;

static int kind_of(std::vector<double> vx, std::vector<double> vy, std::vector<double> vz,
                   double cx, double cy, double cz, double bx, double by, double bz, int* flat64,
                   double* n)
{
    disco_line_loop lp;
    lp.nv = (int)vx.size(); lp.vx = vx.data(); lp.vy = vy.data(); lp.vz = vz.data();
    lp.c[0] = cx; lp.c[1] = cy; lp.c[2] = cz; lp.b[0] = bx; lp.b[1] = by; lp.b[2] = bz;
    LinePack pk;
    const std::string e = pack_line(&lp, 1, 100.0, 100.0, 100.0, pk);
    if (!e.empty() || pk.loops.size() != 1) return -1;
    *flat64 = pk.loops[0].flat64;
    n[0] = pk.loops[0].nx; n[1] = pk.loops[0].ny; n[2] = pk.loops[0].nz;
    return pk.loops[0].kind;
}

int main()
{
    const double PI = 3.14159265358979323846;
    int f64 = -1; double n[3] = {0, 0, 0};
    std::vector<double> x, y, z;
    auto reset = [&]() { x.clear(); y.clear(); z.clear(); };
    // 1 convex hexagon about its centre, b in the plane: glide, frame walk
    reset();
    for (int k = 0; k < 6; k++) { x.push_back(10 * cos(k * PI / 3)); y.push_back(10 * sin(k * PI / 3)); z.push_back(0); }
    { const int kd = kind_of(x, y, z, 0, 0, 0, 1, 0, 0, &f64, n); printf("hexagon %d %d\n", kd, f64); }
    // 2 the same with b along n: prismatic
    { const int kd = kind_of(x, y, z, 0, 0, 0, 0, 0, 2.5, &f64, n); printf("prism %d %d\n", kd, f64); }
    // 3 a concave but simple polygon with the apex outside it (an L shape)
    reset();
    double L[6][2] = {{0,0},{20,0},{20,5},{5,5},{5,20},{0,20}};
    for (auto& p : L) { x.push_back(p[0]); y.push_back(p[1]); z.push_back(0); }
    { const int kd = kind_of(x, y, z, 15, 15, 0, 1, 0, 0, &f64, n); printf("lshape %d %d\n", kd, f64); }
    // 4 the pentagram {5/2}: self-crossing, winding 2 at its centre
    reset();
    for (int k = 0; k < 5; k++) { x.push_back(10 * cos(4 * PI * k / 5)); y.push_back(10 * sin(4 * PI * k / 5)); z.push_back(0); }
    { const int kd = kind_of(x, y, z, 0, 0, 0, 1, 0, 0, &f64, n); printf("pentagram %d %d\n", kd, f64); }
    // 5 a bow tie (one crossing, lobes of opposite sense)
    reset();
    double T[4][2] = {{-10,-5},{10,5},{10,-5},{-10,5}};
    for (auto& p : T) { x.push_back(p[0]); y.push_back(p[1]); z.push_back(0); }
    { const int kd = kind_of(x, y, z, 0, 0.5, 0, 1, 0, 0, &f64, n); printf("bowtie %d %d %.3f %.3f %.3f\n", kd, f64, n[0], n[1], n[2]); }
    // 6 a saddle: off its plane
    reset();
    for (int k = 0; k < 24; k++) { const double t = 2 * PI * k / 24; x.push_back(10 * cos(t)); y.push_back(10 * sin(t)); z.push_back(3 * cos(2 * t)); }
    { const int kd = kind_of(x, y, z, 0, 0, 0, 1, 0, 0, &f64, n); printf("saddle %d %d\n", kd, f64); }
    // 7 collinear back and forth: no vector area; a normal normal to the line
    reset();
    double C[3] = {0, 10, 20};
    for (double c : C) { x.push_back(c); y.push_back(2 * c); z.push_back(0); }
    { const int kd = kind_of(x, y, z, 0, 0, 0, 1, 0, 0, &f64, n); printf("collinear %d %d %.3f\n", kd, f64, n[0] * 1 + n[1] * 2 + n[2] * 0); }
    // 8 a triangle is always simple
    reset();
    double R[3][2] = {{0,0},{10,0},{0,10}};
    for (auto& p : R) { x.push_back(p[0]); y.push_back(p[1]); z.push_back(0); }
    { const int kd = kind_of(x, y, z, 30, 30, 0, 0.3, 0.4, 0, &f64, n); printf("triangle %d %d\n", kd, f64); }
    // 9 the simple-polygon test itself on a star-shaped concave polygon about the origin
    std::vector<double> p1, p2;
    for (int k = 0; k < 10; k++) { const double r = (k & 1) ? 4.0 : 10.0, t = 2 * PI * k / 10; p1.push_back(r * cos(t)); p2.push_back(r * sin(t)); }
    printf("star10 %d\n", simple_polygon(p1, p2) ? 1 : 0);
    // 10 the same points traversed twice round (winding 2, star-shaped): not simple
    std::vector<double> q1, q2;
    for (int k = 0; k < 10; k++) { const double t = 4 * PI * k / 10 + 0.1; q1.push_back(10 * cos(t)); q2.push_back(10 * sin(t)); }
    printf("twice %d\n", simple_polygon(q1, q2) ? 1 : 0);
    // 11 frame-form segment constants of the hexagon: edge B - A, 1/L, w3 t
    reset();
    for (int k = 0; k < 6; k++) { x.push_back(10 * cos(k * PI / 3)); y.push_back(10 * sin(k * PI / 3)); z.push_back(0); }
    disco_line_loop lp;
    lp.nv = 6; lp.vx = x.data(); lp.vy = y.data(); lp.vz = z.data();
    lp.c[0] = lp.c[1] = lp.c[2] = 0; lp.b[0] = 1; lp.b[1] = 0; lp.b[2] = 0;
    LinePack pk;
    pack_line(&lp, 1, 100.0, 100.0, 100.0, pk);
    double worst = 0;
    for (int k = 0; k < 6; k++) {
        SegP g; std::memcpy(&g, &pk.segf[k], sizeof(g));
        const int j = (k + 5) % 6;
        const double e1 = (x[k] - x[j]) * pk.loops[0].e1x + (y[k] - y[j]) * pk.loops[0].e1y;
        const double e2 = (x[k] - x[j]) * pk.loops[0].e2x + (y[k] - y[j]) * pk.loops[0].e2y;
        const double len = std::hypot(e1, e2);
        worst = std::fmax(worst, std::fabs(g.e1 - e1) + std::fabs(g.e2 - e2) + std::fabs(g.iL * len - 1.0) +
                          std::fabs(g.wt1 - g.w3 * g.t1) + std::fabs(g.wt2 - g.w3 * g.t2));
    }
    printf("segconst %.3e\n", worst);
    return 0;
}
'''


@pytest.fixture(scope="module")
def out(tmp_path_factory):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no C++ compiler")
    d = tmp_path_factory.mktemp("pack")
    src = d / "p.cpp"
    src.write_text(SRC)
    exe = d / "p"
    subprocess.run([cxx, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "disco_amd", "csrc"), str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {l.split()[0]: l.split()[1:] for l in lines.strip().splitlines()}


LOOP_3D, LOOP_FLAT, LOOP_GLIDE, LOOP_PRISM = 0, 1, 2, 3


def test_simple_flat_loops_take_the_frame_walks(out):
    assert out["hexagon"] == [str(LOOP_GLIDE), "1"]
    assert out["prism"] == [str(LOOP_PRISM), "1"]
    assert out["lshape"] == [str(LOOP_GLIDE), "1"]          # simple, apex outside: still flat
    assert out["triangle"] == [str(LOOP_GLIDE), "1"]


def test_other_contours_take_the_general_walk(out):
    assert out["pentagram"] == [str(LOOP_3D), "0"]
    assert out["bowtie"][:2] == [str(LOOP_3D), "0"]
    assert out["saddle"] == [str(LOOP_3D), "0"]
    # a collinear back-and-forth contour winds 0 times round every point: the
    # principal argument is exact, so it keeps a frame walk (in the plane of the
    # fallback normal; b has parts in and out of it)
    assert out["collinear"][:2] == [str(LOOP_FLAT), "1"]


def test_normal_of_contours_without_vector_area(out):
    n = np.array([float(v) for v in out["bowtie"][2:]])
    assert abs(np.linalg.norm(n) - 1.0) < 1e-12 and abs(abs(n[2]) - 1.0) < 1e-12
    assert abs(float(out["collinear"][2])) < 1e-12           # normal to the line


def test_simple_polygon_test(out):
    assert out["star10"] == ["1"]
    assert out["twice"] == ["0"]


def test_frame_segment_constants(out):
    assert float(out["segconst"][0]) < 1e-5
