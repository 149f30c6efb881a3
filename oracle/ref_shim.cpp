// ref_shim.cpp -- C entry points onto the reference's own hot-path sources.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's Code/backend_cpu.cpp and Code/loop.cpp, where they lie, into
// oracle/_ref/libdisco_ref.so.  Nothing of the reference is copied: this shim
// only includes its headers and calls its functions, so the golden vectors and
// the oracle's pins come from the reference's compiled code itself.
//
//   ref_barnett_loop      -> barnettLoop<double>   (Code/barnett.hpp:44-107)
//   ref_barnett_loop_f    -> barnettLoop<float>    (Code/barnett.hpp:44-107)
//   ref_kroupa_pair       -> kroupaPair            (Code/kroupa.hpp:75-80)
//   ref_kroupa_cpu_line   -> kroupaCPU, line loops (Code/backend_cpu.cpp:28-63)
//   ref_kroupa_cpu_loopon -> kroupaCPU, loopons    (Code/backend_cpu.cpp:64-99)
//   ref_loop_*            -> class loop, regularPolygon, snapToMidPlane
//                            (Code/loop.cpp:64-244,253-316,426-435,536-608,342-374)
//   ref_lattice           -> generateLattice       (Code/lattice.hpp:35-51)
//   ref_prismatic_slab    -> prismaticSlab         (Code/prismatic.cpp:30-74)
//   ref_classify_burgers  -> classifyBurgers       (Code/lattice.hpp:189-212)
//   ref_frank_copy_shift  -> frankCopyShift        (Code/lattice.hpp:239-256)
//   ref_rotate            -> rotateToLab / rotateToCrystal (Code/lattice.hpp:64-78)
#include <cstring>
#include <iostream>
#include <sstream>
#include <vector>

#include "backend.hpp"
#include "barnett.hpp"
#include "loop.hpp"
#include "prismatic.hpp"

using std::vector;

namespace {
vector<double> v3(const double* p) { return { p[0], p[1], p[2] }; }

// Silences the reference's progress lines on stdout while fn runs.
template <typename F> void quiet(F fn)
{
    std::ostringstream sink;
    std::streambuf* old = std::cout.rdbuf(sink.rdbuf());
    fn();
    std::cout.rdbuf(old);
}

// Shape codes for ref_loop_make
enum { SHAPE_CIRCLE = 0, SHAPE_RECT = 1, SHAPE_PARALLELOGRAM = 2, SHAPE_POLYGON = 3,
       SHAPE_POLYGON_VERTICES = 4 };

// an optional direction: all zero = empty (the default axis)
vector<double> opt3(const double* p)
{
    if (p[0] == 0. && p[1] == 0. && p[2] == 0.) return {};
    return v3(p);
}

// Builds a loop; params by shape:
//   circle:        p[0] = radius
//   rect:          p[0] = half_m, p[1] = half_p, p[2] = corner_r, p[3..5] = side
//   parallelogram: p[0] = half1, p[1] = half2, p[2..4] = dir1, p[5..7] = dir2
//   polygon:       p[0] = sides, p[1] = circumradius R, p[2] = corner_r, p[3..5] = axis
//                  (vertices from regularPolygon(sides, R, {}, m, n))
//   polygon by vertices: p[0] = count, p[1] = corner_r, p[2..4] = axis, p[5..] = vertices
loop make_loop(int shape, const double* p, double da, const double* c, const double* m,
               const double* n, double a0)
{
    const vector<double> one = { 1., 1., 1. };
    switch (shape) {
        case SHAPE_CIRCLE:
            return loop(da, p[0], v3(c), v3(m), v3(n), a0, one);
        case SHAPE_RECT:
            return loop(da, p[0], p[1], p[2], v3(c), v3(m), v3(n), a0, one, opt3(p + 3));
        case SHAPE_PARALLELOGRAM:
            return loop(da, p[0], p[1], v3(p + 2), v3(p + 5), v3(c), v3(m), v3(n), a0, one);
        case SHAPE_POLYGON:
            return loop(da, regularPolygon((int)p[0], p[1], {}, v3(m), v3(n)), p[2], v3(c),
                        v3(m), v3(n), a0, one, opt3(p + 3));
        default: {
            vector<vector<double>> V;
            for (int k = 0; k < (int)p[0]; k++) V.push_back(v3(p + 5 + 3 * k));
            return loop(da, V, p[1], v3(c), v3(m), v3(n), a0, one, opt3(p + 2));
        }
    }
}

void build_line_loops(int nloops, const int* voff, const double* vx, const double* vy,
                      const double* vz, const double* c, const double* phys,
                      vector<LoopData>& loops)
{
    loops.resize(nloops);
    for (int l = 0; l < nloops; l++) {
        LoopData& d = loops[l];
        const double* ph = phys + 10 * l;
        d.phys = { ph[0], ph[1], ph[2], ph[3], ph[4], ph[5], ph[6], ph[7], ph[8], ph[9] };
        d.line = true;
        d.vx.assign(vx + voff[l], vx + voff[l + 1]);
        d.vy.assign(vy + voff[l], vy + voff[l + 1]);
        d.vz.assign(vz + voff[l], vz + voff[l + 1]);
        d.cx = c[3 * l]; d.cy = c[3 * l + 1]; d.cz = c[3 * l + 2];
    }
}
}  // namespace

extern "C" {

void ref_barnett_loop(double px, double py, double pz, const double* vx, const double* vy,
                      const double* vz, int nv, double cx, double cy, double cz, double bx,
                      double by, double bz, double* u)
{
    barnettLoop<double>(px, py, pz, vx, vy, vz, nv, cx, cy, cz, bx, by, bz, u[0], u[1], u[2]);
}

void ref_barnett_loop_f(float px, float py, float pz, const float* vx, const float* vy,
                        const float* vz, int nv, float cx, float cy, float cz, float bx,
                        float by, float bz, float* u)
{
    barnettLoop<float>(px, py, pz, vx, vy, vz, nv, cx, cy, cz, bx, by, bz, u[0], u[1], u[2]);
}

void ref_kroupa_pair(double rx, double ry, double rz, const double* ph, double k1, double b_n,
                     double* u)
{
    const LoopPhys L = { ph[0], ph[1], ph[2], ph[3], ph[4], ph[5], ph[6], ph[7], ph[8], ph[9] };
    kroupaPair(rx, ry, rz, L, k1, b_n, u[0], u[1], u[2]);
}

double ref_kroupa_k1() { return kroupaK1(); }

// kroupaCPU on line loops; phys[10 l ..] is the loop's LoopPhys (b = phys.b * phys.m)
void ref_kroupa_cpu_line(int n, const double* x, const double* y, const double* z, double* ux,
                         double* uy, double* uz, int nloops, const int* voff, const double* vx,
                         const double* vy, const double* vz, const double* c,
                         const double* phys, double Lx, double Ly, double Lz,
                         double* loop_sec)
{
    vector<LoopData> loops;
    build_line_loops(nloops, voff, vx, vy, vz, c, phys, loops);
    const AtomArrays atoms = { x, y, z, ux, uy, uz, n };
    BackendTiming t;
    quiet([&] { kroupaCPU(atoms, loops, Lx, Ly, Lz, t); });
    if (loop_sec)
        for (int l = 0; l < nloops; l++) loop_sec[l] = t.loopSec[l];
}

void ref_kroupa_cpu_loopon(int n, const double* x, const double* y, const double* z,
                           double* ux, double* uy, double* uz, int nloops, const int* poff,
                           const double* px, const double* py, const double* pz,
                           const double* phys, double Lx, double Ly, double Lz,
                           double* loop_sec)
{
    vector<LoopData> loops(nloops);
    for (int l = 0; l < nloops; l++) {
        LoopData& d = loops[l];
        const double* ph = phys + 10 * l;
        d.phys = { ph[0], ph[1], ph[2], ph[3], ph[4], ph[5], ph[6], ph[7], ph[8], ph[9] };
        d.x.assign(px + poff[l], px + poff[l + 1]);
        d.y.assign(py + poff[l], py + poff[l + 1]);
        d.z.assign(pz + poff[l], pz + poff[l + 1]);
    }
    const AtomArrays atoms = { x, y, z, ux, uy, uz, n };
    BackendTiming t;
    quiet([&] { kroupaCPU(atoms, loops, Lx, Ly, Lz, t); });
    if (loop_sec)
        for (int l = 0; l < nloops; l++) loop_sec[l] = t.loopSec[l];
}

// Polyline of a loop (Code/loop.cpp:536-580) at tolerance eps.  Returns the
// vertex count; writes up to cap vertices into out (x0 y0 z0 x1 ...).
int ref_loop_polyline(int shape, const double* p, const double* c, const double* m,
                      const double* n, double a0, double eps, double* out, int cap)
{
    const loop lp = make_loop(shape, p, 1.0, c, m, n, a0);
    const vector<vector<double>> P = lp.polyline(eps);
    for (int k = 0; k < (int)P.size() && k < cap; k++)
        for (int i = 0; i < 3; i++) out[3 * k + i] = P[k][i];
    return (int)P.size();
}

// Loop physics as the driver assembles it (Code/main.cpp:698-699):
// phys = {|b| x3, m_hat, n_hat, dS} after generateSurface(); area = contour area.
void ref_loop_phys(int shape, const double* p, double da, const double* c, const double* m,
                   const double* n, double a0, double* phys, double* area)
{
    loop lp = make_loop(shape, p, da, c, m, n, a0);
    lp.generateSurface();
    const vector<double> b = lp.getBurgers(), mh = lp.getM(), nh = lp.getN();
    const double out[10] = { b[0], b[1], b[2], mh[0], mh[1], mh[2], nh[0], nh[1], nh[2],
                             lp.getSurface() };
    std::memcpy(phys, out, sizeof(out));
    *area = lp.getContourArea();
}

// Loopon grid (Code/loop.cpp:253-273).  Returns the count; writes up to cap.
int ref_loop_loopons(int shape, const double* p, double da, const double* c, const double* m,
                     const double* n, double a0, double* out, int cap)
{
    loop lp = make_loop(shape, p, da, c, m, n, a0);
    lp.generateSurface();
    lp.generateLoopons();
    const int np = lp.getNumberOfLoopons();
    for (int k = 0; k < np && k < cap; k++) {
        const vector<double> q = lp.getOneLoopon(k);
        for (int i = 0; i < 3; i++) out[3 * k + i] = q[i];
    }
    return np;
}

// regularPolygon (Code/loop.cpp:585-608), default vertex axis
void ref_regular_polygon(int sides, double R, const double* m, const double* n, double* out)
{
    const vector<vector<double>> P = regularPolygon(sides, R, {}, v3(m), v3(n));
    for (int k = 0; k < sides; k++)
        for (int i = 0; i < 3; i++) out[3 * k + i] = P[k][i];
}

// snapToMidPlane (Code/loop.cpp:342-374); lattice 0 = BCC, 1 = FCC
void ref_snap_mid_plane(const double* center, const double* n, double a0, int nx, int ny,
                        int nz, int lattice, double* out, double* spacing)
{
    const vector<double> s = snapToMidPlane(v3(center), v3(n), a0, nx, ny, nz, *spacing,
                                            lattice ? Lattice::FCC : Lattice::BCC);
    for (int i = 0; i < 3; i++) out[i] = s[i];
}

// generateLattice (Code/lattice.hpp:35-51) into SoA arrays of capacity cap
int ref_lattice(int lattice, int nx, int ny, int nz, double a0, double* x, double* y, double* z,
                int cap)
{
    const vector<vector<double>> A =
        generateLattice(lattice ? Lattice::FCC : Lattice::BCC, nx, ny, nz, a0);
    for (int i = 0; i < (int)A.size() && i < cap; i++) { x[i] = A[i][0]; y[i] = A[i][1]; z[i] = A[i][2]; }
    return (int)A.size();
}

// prismaticSlab (Code/prismatic.cpp:30-74) of one loop over SoA atoms; returns
// the slab size (atoms / side written up to cap) and info = (type, planes,
// height, spacing, b0, b1, b2), type 0 none, 1 vacancy, 2 interstitial.
int ref_prismatic_slab(int shape, const double* p, const double* c, const double* m,
                       const double* n, double a0, double spacing, int natoms, const double* x,
                       const double* y, const double* z, double lx, double ly, double lz,
                       int* atoms, int* side, int cap, double* info)
{
    const loop lp = make_loop(shape, p, 1.0, c, m, n, a0);
    const vector<double> X(x, x + natoms), Y(y, y + natoms), Z(z, z + natoms);
    const SlabResult r = prismaticSlab(lp, spacing, X, Y, Z, lx, ly, lz);
    const int ns = (int)r.atoms.size();
    for (int k = 0; k < ns && k < cap; k++) { atoms[k] = r.atoms[k]; side[k] = r.side[k]; }
    info[0] = r.type == MatterType::NONE ? 0 : (r.type == MatterType::VACANCY ? 1 : 2);
    info[1] = r.planes;
    info[2] = r.height;
    info[3] = r.spacing;
    for (int i = 0; i < 3; i++) info[4 + i] = r.b[i];
    return ns;
}

// classifyBurgers (Code/lattice.hpp:189-212): 0 perfect, 1 Frank, 2 Shockley, 3 other
int ref_classify_burgers(int lattice, const double* m, const double* n)
{
    switch (classifyBurgers(lattice ? Lattice::FCC : Lattice::BCC, v3(m), v3(n))) {
        case BurgersClass::PERFECT: return 0;
        case BurgersClass::FRANK: return 1;
        case BurgersClass::SHOCKLEY: return 2;
        default: return 3;
    }
}

// frankCopyShift (Code/lattice.hpp:239-256)
void ref_frank_copy_shift(const double* v, const double* nhat, double a0, double* out)
{
    frankCopyShift(v, nhat, a0, out);
}

// rotateToLab (to_lab != 0) or rotateToCrystal of v with the row-major R
void ref_rotate(const double* R, int to_lab, const double* v, double* out)
{
    Orientation o;
    o.active = true;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) o.R[i][j] = R[3 * i + j];
    const vector<double> r = to_lab ? rotateToLab(o, v3(v)) : rotateToCrystal(o, v3(v));
    for (int i = 0; i < 3; i++) out[i] = r[i];
}

}  // extern "C"
