// plan_pack.h -- host packing of loop geometry into the device tables.
//
// Line loops (the reference's LoopData with line = true, Code/backend.hpp:20-26,
// assembled at Code/main.cpp:689-711): vertices become offsets from the apex c
// in FP64; segment k = V[k-1] -> V[k] with V[nv-1] -> V[0] closing the contour
// (Code/barnett.hpp:62-64).  Zero-length segments are dropped: the reference
// skips their line terms (Code/barnett.hpp:77) and their fan triangle has
// num = 0, den >= 0, so atan2 gives 0 -- they contribute nothing.
//
// Loopon loops (line = false): loopons become offsets from their centroid, as
// the reference's --fp32 upload does (Code/backend_cuda.cu:676-685).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "disco_hip.h"
#include "pair_terms.h"

namespace disco {

struct LinePack {
    std::vector<LineLoop> loops;   // ordered by kind: loops [kofs[k], kofs[k+1]) have kind k
    int kofs[5] = {0, 0, 0, 0, 0};
    std::vector<SegF> segf;
    std::vector<SegD> segd;
    double segments = 0;      // total segment count (work per atom-image)
    std::vector<double> work; // per input loop: its segment count
};

struct LooponPack {
    std::vector<LooponLoop> loops;
    std::vector<int> src;          // input index of each packed loop
    std::vector<float4> pf;
    std::vector<double4> pd;
    double loopons = 0;
    std::vector<double> work;
};

// A normal for a contour with no vector area: the plane of the vertex
// farthest from V[0] and the one that spans the largest parallelogram with it
// (for collinear vertices, any direction normal to their line).
inline void fallback_normal(const std::vector<double>& ax, const std::vector<double>& ay,
                            const std::vector<double>& az, double& Nx, double& Ny, double& Nz)
{
    const int nv = (int)ax.size();
    int kf = 0;
    double best = -1.0;
    for (int k = 0; k < nv; k++) {
        const double d2 = (ax[k] - ax[0]) * (ax[k] - ax[0]) + (ay[k] - ay[0]) * (ay[k] - ay[0]) +
                          (az[k] - az[0]) * (az[k] - az[0]);
        if (d2 > best) { best = d2; kf = k; }
    }
    const double ux = ax[kf] - ax[0], uy = ay[kf] - ay[0], uz = az[kf] - az[0];
    best = 0.0;
    Nx = Ny = Nz = 0.0;
    for (int k = 0; k < nv; k++) {
        const double wx = ax[k] - ax[0], wy = ay[k] - ay[0], wz = az[k] - az[0];
        const double cx = uy * wz - uz * wy, cy = uz * wx - ux * wz, cz = ux * wy - uy * wx;
        const double c2 = cx * cx + cy * cy + cz * cz;
        if (c2 > best) { best = c2; Nx = cx; Ny = cy; Nz = cz; }
    }
    if (!(best > 1e-24 * (ux * ux + uy * uy + uz * uz) * (ux * ux + uy * uy + uz * uz))) {
        // collinear: u x (the axis along u's smallest component)
        const double fx = std::fabs(ux), fy = std::fabs(uy), fz = std::fabs(uz);
        const double ex = (fx <= fy && fx <= fz) ? 1.0 : 0.0;
        const double ey = (ex == 0.0 && fy <= fz) ? 1.0 : 0.0;
        const double ez = (ex == 0.0 && ey == 0.0) ? 1.0 : 0.0;
        Nx = uy * ez - uz * ey; Ny = uz * ex - ux * ez; Nz = ux * ey - uy * ex;
    }
}

// True when the closed polygon (p1[k], p2[k]) in its plane is simple, so that
// its winding number is 0 or +-1 about every point: star-shaped about the
// origin (the apex: every fan triangle turns the same way, once round), or
// else no two non-adjacent edges touch.  Ties count as touching, so a
// degenerate polygon is called not simple (and summed by the general walk).
inline bool simple_polygon(const std::vector<double>& p1, const std::vector<double>& p2)
{
    const int n = (int)p1.size();
    if (n <= 3) return true;
    bool pos = true, neg = true;
    double turn = 0.0;
    for (int k = 0; k < n; k++) {
        const int j = (k + n - 1) % n;
        const double cr = p1[j] * p2[k] - p2[j] * p1[k];
        pos = pos && cr > 0.0;
        neg = neg && cr < 0.0;
        turn += std::atan2(cr, p1[j] * p1[k] + p2[j] * p2[k]);
    }
    if ((pos || neg) && std::fabs(turn) < 3.0 * 3.14159265358979323846) return true;
    auto orient = [&](int a, int b, int c) {
        return (p1[b] - p1[a]) * (p2[c] - p2[a]) - (p2[b] - p2[a]) * (p1[c] - p1[a]);
    };
    auto straddle = [](double u, double v) { return (u <= 0.0 && v >= 0.0) || (u >= 0.0 && v <= 0.0); };
    // edge k runs V[k-1] -> V[k]
    for (int a = 0; a < n; a++)
        for (int b = a + 2; b < n; b++) {
            if (a == 0 && b == n - 1) continue;   // adjacent through V[n-1] -> V[0]
            const int a0 = (a + n - 1) % n, b0 = b - 1;
            if (straddle(orient(a0, a, b0), orient(a0, a, b)) &&
                straddle(orient(b0, b, a0), orient(b0, b, a)))
                return false;
        }
    return true;
}

// Height of the near slab: a lane whose loop image lies within this distance
// of the loop's plane (and inside its radius + this margin) is evaluated in
// FP64 in the FP32C mode.
inline double near_height(double bmag) { return std::fmax(3.0 * bmag, 6.0); }

// The 9-point cubature of a flat cut (the fields rw0 ... rR of LineLoop):
// p1, p2 are the vertices relative to the apex in the loop frame (e1 along
// the farthest vertex).  Its centre weight w0 and eight ring points of weight
// w1 at radius rho (angles j pi/4 from e1) match the cut's area A,
// I2 = int r^2 dA and I4 = int r^4 dA: w0 + 8 w1 = A, 8 w1 rho^2 = I2,
// 8 w1 rho^4 = I4 (for a disc: rho^2 = 2R^2/3, w0 = A/4, w1 = 3A/32).  The
// ring of eight is isotropic through rank 7, so the rule's first angular
// error is at degree 8 and its degree-6 error is the radial I6 alone.  The
// rule is used (rule_ok) only if it also matches every other moment
// int x^a y^b dA, a + b <= 5, of the cut to 1e-7 A R^(a+b): then the far
// field's error is O((R/r)^6) of the field (cuts with moments isotropic
// through rank 5 -- circles, regular polygons of six sides or more).  The
// cut's moments come from its fan triangles (c, V[k-1], V[k]) with Radon's
// 7-point rule on a triangle, which is exact for polynomials of degree 5.
inline void pack_rule9(const std::vector<double>& p1, const std::vector<double>& p2,
                       const double b[3], const double e1[3], const double e2[3],
                       const double n[3], double rmax, LineLoop& L)
{
    const int nv = (int)p1.size();
    double M[6][6] = {};   // M[a][b] = int x^a y^b dA, a + b <= 5
    const double s15 = std::sqrt(15.0);
    const double ta = (6.0 - s15) / 21.0, tb = (9.0 + 2.0 * s15) / 21.0;
    const double tc = (6.0 + s15) / 21.0, td = (9.0 - 2.0 * s15) / 21.0;
    const double wa = (155.0 - s15) / 1200.0, wc = (155.0 + s15) / 1200.0;
    const double bary[7][3] = {{1.0 / 3, 1.0 / 3, 1.0 / 3}, {ta, ta, tb}, {ta, tb, ta},
                               {tb, ta, ta}, {tc, tc, td}, {tc, td, tc}, {td, tc, tc}};
    const double bw[7] = {9.0 / 40.0, wa, wa, wa, wc, wc, wc};
    for (int k = 0; k < nv; k++) {
        const int j = (k + nv - 1) % nv;
        const double area = 0.5 * (p1[j] * p2[k] - p2[j] * p1[k]);   // signed, apex at 0
        for (int q = 0; q < 7; q++) {
            const double x = bary[q][1] * p1[j] + bary[q][2] * p1[k];
            const double y = bary[q][1] * p2[j] + bary[q][2] * p2[k];
            double xa = 1.0;
            for (int a = 0; a <= 5; a++) {
                double yb = 1.0;
                for (int c = 0; a + c <= 5; c++) {
                    M[a][c] += area * bw[q] * xa * yb;
                    yb *= y;
                }
                xa *= x;
            }
        }
    }
    const double A = M[0][0], I2 = M[2][0] + M[0][2], I4 = M[4][0] + 2.0 * M[2][2] + M[0][4];
    L.rule_ok = 0;
    L.rR = (float)rmax;
    if (!(A > 0.0 && I2 > 0.0 && I4 > 0.0 && rmax > 0.0)) return;
    const double rho = std::sqrt(I4 / I2), w1 = I2 * I2 / (8.0 * I4), w0 = A - 8.0 * w1;
    if (!(w0 >= 0.0)) return;
    // the rule's own moments
    double R[6][6] = {};
    R[0][0] = w0;
    for (int j = 0; j < 8; j++) {
        const double th = j * (kPi / 4.0);
        const double x = rho * std::cos(th), y = rho * std::sin(th);
        double xa = 1.0;
        for (int a = 0; a <= 5; a++) {
            double yb = 1.0;
            for (int c = 0; a + c <= 5; c++) {
                R[a][c] += w1 * xa * yb;
                yb *= y;
            }
            xa *= x;
        }
    }
    double worst = 0.0, rp = 1.0;
    for (int d = 0; d <= 5; d++, rp *= rmax)
        for (int a = 0; a <= d; a++)
            worst = std::fmax(worst, std::fabs(M[a][d - a] - R[a][d - a]) / (A * rp));
    const double bmag = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    if (!(worst <= 1e-7) || !(bmag > 0.0)) return;
    const double mf[3] = {(b[0] * e1[0] + b[1] * e1[1] + b[2] * e1[2]) / bmag,
                          (b[0] * e2[0] + b[1] * e2[1] + b[2] * e2[2]) / bmag,
                          (b[0] * n[0] + b[1] * n[1] + b[2] * n[2]) / bmag};
    const double c1 = -kC3 * (1.0 - 2.0 * kNu), c2 = -3.0 * kC3;   // per unit area
    L.rw0 = (float)w0;
    L.rw1 = (float)w1;
    L.rrho = (float)rho;
    for (int i = 0; i < 3; i++) {
        L.rm[i] = (float)(c2 * mf[i]);
        L.rmf[i] = (float)mf[i];
    }
    const double s2 = std::sqrt(0.5);
    L.rmp[0] = (float)(c2 * mf[0] * rho);
    L.rmp[1] = (float)(c2 * mf[1] * rho);
    L.rmp[2] = (float)(c2 * (mf[0] + mf[1]) * s2 * rho);
    L.rmp[3] = (float)(c2 * (mf[1] - mf[0]) * s2 * rho);
    L.rcb = (float)(c1 * mf[2]);
    L.rbm = (float)bmag;
    L.rt0 = 0.0f;
    L.rwinv = 0.0f;
    L.rule_ok = 1;
}

// Two-ring cubature of a flat cut for lanes between ~3 and ~5 cut radii of
// the loop image (the fields bw0 ... bmp2 of LineLoop; call after pack_rule9
// succeeded, whose m, c1 and c2 fields it shares).  The centre and two rings
// of twelve points are a Gauss-Radau rule in s = r^2 for the cut's radial
// moments: the ring radii^2 s1 < s2 are the roots of the monic quadratic
// orthogonal to 1 and s under s dmu(s) (dmu: the cut's area by s), so that
// w0 + 12 (w1 + w2) = I0 and 12 (w1 s1^k + w2 s2^k) = I_2k for k = 1..4 (for a
// disc: s = (6 -+ sqrt 6) R^2 / 10, centre A / 9).  A ring of twelve is
// isotropic through rank 11.  The rule is used (band_ok) only if it matches
// every moment int x^a y^b dA, a + b <= 9, of the cut to 1e-7 A R^(a+b), and
// the cut's moments of degrees 10 and 11 are isotropic to the same bound:
// then its first error terms are the radial I10 mismatch and the angular
// terms of degree 12, O((R/r)^12) of the field (for a 48-gon and
// |b| = 2.56 Angstrom at most 1.2e-7 Angstrom at r = 3R).  The moments are
// exact: over each fan triangle (0, A, B), x^a y^b in the barycentric
// coordinates (u, v) of P = u A + v B is a polynomial whose terms integrate
// as p! q! / (p + q + 2)!.
inline void pack_rule_band(const std::vector<double>& p1, const std::vector<double>& p2,
                           double rmax, LineLoop& L)
{
    constexpr int D = 11;
    L.band_ok = 0;
    L.rtb = 3.0e38f;
    if (!L.rule_ok) return;
    const int nv = (int)p1.size();
    double fact[2 * D + 3];
    fact[0] = 1.0;
    for (int i = 1; i < 2 * D + 3; i++) fact[i] = fact[i - 1] * i;
    auto binom = [&](int n, int k) { return fact[n] / (fact[k] * fact[n - k]); };
    double M[D + 1][D + 1] = {};   // M[a][b] = int x^a y^b dA, a + b <= D
    for (int k = 0; k < nv; k++) {
        const int j = (k + nv - 1) % nv;
        const double A1 = p1[j], A2 = p2[j], B1 = p1[k], B2 = p2[k];
        const double twice_area = A1 * B2 - A2 * B1;                   // signed
        double pa1[D + 1], pb1[D + 1], pa2[D + 1], pb2[D + 1];
        pa1[0] = pb1[0] = pa2[0] = pb2[0] = 1.0;
        for (int i = 1; i <= D; i++) {
            pa1[i] = pa1[i - 1] * A1; pb1[i] = pb1[i - 1] * B1;
            pa2[i] = pa2[i - 1] * A2; pb2[i] = pb2[i - 1] * B2;
        }
        for (int a = 0; a <= D; a++)
            for (int b = 0; a + b <= D; b++) {
                double sum = 0.0;
                for (int i = 0; i <= a; i++)          // u^i v^(a-i) from x^a
                    for (int jj = 0; jj <= b; jj++) { // u^jj v^(b-jj) from y^b
                        const int pu = i + jj, pv = a + b - pu;
                        sum += binom(a, i) * binom(b, jj) * pa1[i] * pb1[a - i] * pa2[jj] *
                               pb2[b - jj] * fact[pu] * fact[pv] / fact[a + b + 2];
                    }
                M[a][b] += twice_area * sum;
            }
    }
    auto radial = [&](int k) {         // I_2k = int (x^2 + y^2)^k dA
        double s = 0.0;
        for (int j = 0; j <= k; j++) s += binom(k, j) * M[2 * j][2 * k - 2 * j];
        return s;
    };
    const double I0 = radial(0), I2 = radial(1), I4 = radial(2), I6 = radial(3), I8 = radial(4);
    // p(s) = s^2 + c1 s + c0 with int p s^j s dmu = 0, j = 0, 1
    const double det = I4 * I4 - I2 * I6;
    if (!(std::fabs(det) > 0.0)) return;
    const double c1 = (-I6 * I4 + I8 * I2) / det, c0 = (-I8 * I4 + I6 * I6) / det;
    const double disc = c1 * c1 - 4.0 * c0;
    if (!(disc > 0.0)) return;
    const double sq = std::sqrt(disc);
    const double s1 = 0.5 * (-c1 - sq), s2 = 0.5 * (-c1 + sq);
    if (!(s1 > 0.0 && s2 > s1)) return;
    const double W1 = (I2 * s2 - I4) / (s2 - s1), W2 = (I4 - I2 * s1) / (s2 - s1);
    const double t1w = W1 / s1, t2w = W2 / s2, w0 = I0 - t1w - t2w;   // ring totals, centre
    if (!(t1w > 0.0 && t2w > 0.0 && w0 >= 0.0)) return;
    const double r1 = std::sqrt(s1), r2 = std::sqrt(s2), w1 = t1w / 12.0, w2 = t2w / 12.0;
    // the rule's own moments, and the check
    double Rm[D + 1][D + 1] = {};
    Rm[0][0] = w0;
    for (int ring = 0; ring < 2; ring++)
        for (int j = 0; j < 12; j++) {
            const double th = j * (kPi / 6.0) + (ring ? kPi / 12.0 : 0.0);
            const double rr = ring ? r2 : r1, ww = ring ? w2 : w1;
            const double x = rr * std::cos(th), y = rr * std::sin(th);
            double xa = 1.0;
            for (int a = 0; a <= D; a++) {
                double yb = 1.0;
                for (int b = 0; a + b <= D; b++) {
                    Rm[a][b] += ww * xa * yb;
                    yb *= y;
                }
                xa *= x;
            }
        }
    const double A = I0;
    double worst = 0.0, rp = 1.0;
    for (int d = 0; d <= D; d++, rp *= rmax) {
        const double Id = d % 2 == 0 ? radial(d / 2) : 0.0;
        for (int a = 0; a <= d; a++) {
            const int b = d - a;
            double want;
            if (d <= 9) {
                want = Rm[a][b];
            } else {       // isotropic: I_d (a-1)!! (b-1)!! / d!! for even a, b
                want = 0.0;
                if (a % 2 == 0 && b % 2 == 0) {
                    double num = 1.0, den = 1.0;
                    for (int t = a - 1; t > 0; t -= 2) num *= t;
                    for (int t = b - 1; t > 0; t -= 2) num *= t;
                    for (int t = d; t > 0; t -= 2) den *= t;
                    want = Id * num / den;
                }
            }
            worst = std::fmax(worst, std::fabs(M[a][b] - want) / (A * rp));
        }
    }
    if (!(worst <= 1e-7)) return;
    L.bw0 = (float)w0;
    L.bw1 = (float)w1;
    L.bw2 = (float)w2;
    L.brho1 = (float)r1;
    L.brho2 = (float)r2;
    // rm = c2 m in the frame (pack_rule9): c2 m . P = rho (rm1 cos th + rm2 sin th)
    for (int j = 0; j < 6; j++) {
        const double th1 = j * (kPi / 6.0), th2 = th1 + kPi / 12.0;
        L.bmp1[j] = (float)(r1 * ((double)L.rm[0] * std::cos(th1) + (double)L.rm[1] * std::sin(th1)));
        L.bmp2[j] = (float)(r2 * ((double)L.rm[0] * std::cos(th2) + (double)L.rm[1] * std::sin(th2)));
    }
    L.band_ok = 1;
}

// Image shifts of every loop of a line plan in the loop's frame, appended to
// the FP64 table as 128-B slots of four entries: entry
// ((i + img) n + (j + img)) n + (k + img), n = 2 img + 1, of loop L holds
// (i X + j Y) + k Z (three doubles and a pad) with the box edges X, Y, Z of its
// record, computed with contraction off -- the expression the kernels would
// otherwise evaluate per lane, so the image coordinates q = Q - shift come out
// bit for bit the same.  L.sh0 is the loop's first slot.  The FP32 table gets
// as many empty slots, so that the two tables keep one index (seg0) per
// segment.  Returns the bytes added to the FP64 table.
inline size_t pack_shifts(LinePack& pk, int img)
{
#pragma clang fp contract(off)
    const int n = 2 * img + 1, nimg = n * n * n, slots = (nimg + 3) / 4;
    std::vector<double> buf((size_t)slots * 16);
    for (LineLoop& L : pk.loops) {
        L.sh0 = (int)pk.segd.size();
        std::fill(buf.begin(), buf.end(), 0.0);
        for (int di = -img; di <= img; di++)
            for (int dj = -img; dj <= img; dj++)
                for (int dk = -img; dk <= img; dk++) {
                    const double i = di, j = dj, k = dk;
                    double* e = buf.data() + 4 * (((size_t)(di + img) * n + (dj + img)) * n +
                                                  (dk + img));
                    e[0] = (i * L.X1 + j * L.Y1) + k * L.Z1;
                    e[1] = (i * L.X2 + j * L.Y2) + k * L.Z2;
                    e[2] = (i * L.X3 + j * L.Y3) + k * L.Z3;
                }
        SegF empty;
        std::memset(&empty, 0, sizeof(empty));
        for (int t = 0; t < slots; t++) {
            SegD d;
            std::memcpy(&d, buf.data() + 16 * (size_t)t, sizeof(d));
            pk.segd.push_back(d);
            pk.segf.push_back(empty);
        }
    }
    return pk.loops.size() * (size_t)slots * sizeof(SegD);
}

// Returns "" on success, else the reason the input is rejected.
inline std::string pack_line(const disco_line_loop* in, int nloops, double Lx, double Ly,
                             double Lz, LinePack& out)
{
    out = LinePack();
    out.work.assign(nloops, 0.0);
    for (int l = 0; l < nloops; l++) {
        const disco_line_loop& lp = in[l];
        if (lp.nv < 0) return "negative vertex count";
        if (lp.nv > 0 && (!lp.vx || !lp.vy || !lp.vz)) return "null vertex array";
        // distinct consecutive vertices, relative to the apex
        std::vector<double> ax, ay, az;
        for (int k = 0; k < lp.nv; k++) {
            const double x = lp.vx[k] - lp.c[0], y = lp.vy[k] - lp.c[1], z = lp.vz[k] - lp.c[2];
            if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z))
                return "non-finite vertex";
            if (!ax.empty() && x == ax.back() && y == ay.back() && z == az.back()) continue;
            ax.push_back(x); ay.push_back(y); az.push_back(z);
        }
        while (ax.size() > 1 && ax.back() == ax[0] && ay.back() == ay[0] && az.back() == az[0]) {
            ax.pop_back(); ay.pop_back(); az.pop_back();
        }
        const int nv = (int)ax.size();
        out.work[l] = (double)lp.nv;
        if (nv < 3) continue;   // degenerate contour: its field is identically zero
        // vector area -> plane normal
        double Nx = 0, Ny = 0, Nz = 0;
        for (int k = 0; k < nv; k++) {
            const int j = (k + nv - 1) % nv;
            Nx += ay[j] * az[k] - az[j] * ay[k];
            Ny += az[j] * ax[k] - ax[j] * az[k];
            Nz += ax[j] * ay[k] - ay[j] * ax[k];
        }
        double Nn = std::sqrt(Nx * Nx + Ny * Ny + Nz * Nz);
        double vmax2 = 0.0;
        for (int k = 0; k < nv; k++)
            vmax2 = std::fmax(vmax2, (ax[k] - ax[0]) * (ax[k] - ax[0]) +
                                         (ay[k] - ay[0]) * (ay[k] - ay[0]) +
                                         (az[k] - az[0]) * (az[k] - az[0]));
        if (!(Nn > 1e-12 * vmax2)) {
            // no plane from the vector area (a contour whose lobes cancel):
            // any normal serves, the loop is summed in global axes
            fallback_normal(ax, ay, az, Nx, Ny, Nz);
            Nn = std::sqrt(Nx * Nx + Ny * Ny + Nz * Nz);
        }
        const double nx = Nx / Nn, ny = Ny / Nn, nz = Nz / Nn;
        double hmax = 0.0, rmax = 0.0;
        for (int k = 0; k < nv; k++) {
            const double h = ax[k] * nx + ay[k] * ny + az[k] * nz;
            const double r2 = ax[k] * ax[k] + ay[k] * ay[k] + az[k] * az[k] - h * h;
            hmax = std::fmax(hmax, std::fabs(h));
            rmax = std::fmax(rmax, std::sqrt(std::fmax(r2, 0.0)));
        }

        LineLoop L;
        std::memset(&L, 0, sizeof(L));
        L.cx = lp.c[0]; L.cy = lp.c[1]; L.cz = lp.c[2];
        L.bx = lp.b[0]; L.by = lp.b[1]; L.bz = lp.b[2];
        L.nx = nx; L.ny = ny; L.nz = nz;
        const double bmag = std::sqrt(lp.b[0] * lp.b[0] + lp.b[1] * lp.b[1] + lp.b[2] * lp.b[2]);
        const double H = near_height(bmag);
        L.hn = H + hmax;
        L.rn2 = (rmax + H) * (rmax + H);
        L.seg0 = (int)out.segd.size();
        L.nv = nv;
        L.src = l;

        // loop frame: e1 along the in-plane offset of the vertex farthest from
        // the apex, e2 = n x e1
        int kf = 0;
        double best = -1.0;
        for (int k = 0; k < nv; k++) {
            const double h = ax[k] * nx + ay[k] * ny + az[k] * nz;
            const double r2 = ax[k] * ax[k] + ay[k] * ay[k] + az[k] * az[k] - h * h;
            if (r2 > best) { best = r2; kf = k; }
        }
        {
            const double h = ax[kf] * nx + ay[kf] * ny + az[kf] * nz;
            double ex = ax[kf] - h * nx, ey = ay[kf] - h * ny, ez = az[kf] - h * nz;
            const double en = std::sqrt(ex * ex + ey * ey + ez * ez);
            L.e1x = ex / en; L.e1y = ey / en; L.e1z = ez / en;
            L.e2x = ny * L.e1z - nz * L.e1y;
            L.e2y = nz * L.e1x - nx * L.e1z;
            L.e2z = nx * L.e1y - ny * L.e1x;
        }
        // FP32 layout: the frame form when the vertices sit in the plane to
        // float resolution (their heights are then below the rounding of their
        // in-plane coordinates) and every projected segment has a length
        const double res = std::ldexp(std::fmax(rmax, 1.0), -24);
        std::vector<double> p1(nv), p2(nv);
        for (int k = 0; k < nv; k++) {
            p1[k] = ax[k] * L.e1x + ay[k] * L.e1y + az[k] * L.e1z;
            p2[k] = ax[k] * L.e2x + ay[k] * L.e2y + az[k] * L.e2z;
        }
        // the frame forms take the solid angle's principal argument, exact
        // for a simple polygon only
        const bool simple = simple_polygon(p1, p2);
        bool flat = hmax <= res && simple;
        for (int k = 0; k < nv && flat; k++) {
            const int j = (k + nv - 1) % nv;
            flat = std::hypot(p1[k] - p1[j], p2[k] - p2[j]) > res;
        }
        // FP64 layout: the frame form (SegPD) only when the vertices sit in the
        // plane to double resolution (every DISCO shape: heights ~1e-16 of its
        // size); otherwise Seg<double> in global axes
        const double res64 = std::ldexp(std::fmax(rmax, 1.0), -48);
        bool flat64 = hmax <= res64 && simple;
        for (int k = 0; k < nv && flat64; k++) {
            const int j = (k + nv - 1) % nv;
            flat64 = std::hypot(p1[k] - p1[j], p2[k] - p2[j]) > res64;
        }
        L.flat64 = flat64 ? 1 : 0;
        const double b1 = lp.b[0] * L.e1x + lp.b[1] * L.e1y + lp.b[2] * L.e1z;
        const double b2 = lp.b[0] * L.e2x + lp.b[1] * L.e2y + lp.b[2] * L.e2z;
        const double b3 = lp.b[0] * nx + lp.b[1] * ny + lp.b[2] * nz;
        const double btiny = std::ldexp(bmag, -24);
        L.kind = !flat ? LOOP_3D
               : std::fabs(b3) <= btiny ? LOOP_GLIDE
               : std::hypot(b1, b2) <= btiny ? LOOP_PRISM : LOOP_FLAT;
        L.reach = std::sqrt(L.rn2 + L.hn * L.hn) * (1.0 + 1e-12) + 1e-6;
        L.iLx = Lx > 0 ? 1.0 / Lx : 0.0;
        L.iLy = Ly > 0 ? 1.0 / Ly : 0.0;
        L.iLz = Lz > 0 ? 1.0 / Lz : 0.0;
        // box edges in the frame; float copies for the FP32 image fields
        L.X1 = Lx * L.e1x; L.X2 = Lx * L.e2x; L.X3 = Lx * nx;
        L.Y1 = Ly * L.e1y; L.Y2 = Ly * L.e2y; L.Y3 = Ly * ny;
        L.Z1 = Lz * L.e1z; L.Z2 = Lz * L.e2z; L.Z3 = Lz * nz;
        const double e1v[3] = {L.e1x, L.e1y, L.e1z}, e2v[3] = {L.e2x, L.e2y, L.e2z};
        const double nv3[3] = {nx, ny, nz};
        for (int i = 0; i < 3; i++) {
            L.bf[i] = (float)lp.b[i];
            L.e1f[i] = (float)e1v[i];
            L.e2f[i] = (float)e2v[i];
            L.nf[i] = (float)nv3[i];
        }
        if (L.kind != LOOP_3D) {
            const double n3[3] = {nx, ny, nz};
            pack_rule9(p1, p2, lp.b, e1v, e2v, n3, rmax, L);
        }
        pack_rule_band(p1, p2, rmax, L);
        const double fb3 = L.kind == LOOP_GLIDE ? 0.0 : b3;
        const double fb1 = L.kind == LOOP_PRISM ? 0.0 : b1;
        const double fb2 = L.kind == LOOP_PRISM ? 0.0 : b2;

        for (int k = 0; k < nv; k++) {
            const int j = (k + nv - 1) % nv;
            SegD d;
            std::memset(&d, 0, sizeof(d));
            d.bx = ax[k]; d.by = ay[k]; d.bz = az[k];
            const double ex = ax[k] - ax[j], ey = ay[k] - ay[j], ez = az[k] - az[j];
            const double len = std::sqrt(ex * ex + ey * ey + ez * ez);
            d.L = len;
            d.tx = ex / len; d.ty = ey / len; d.tz = ez / len;
            d.nx = ay[j] * az[k] - az[j] * ay[k];
            d.ny = az[j] * ax[k] - ax[j] * az[k];
            d.nz = ax[j] * ay[k] - ay[j] * ax[k];
            const double wx = lp.b[1] * d.tz - lp.b[2] * d.ty;
            const double wy = lp.b[2] * d.tx - lp.b[0] * d.tz;
            const double wz = lp.b[0] * d.ty - lp.b[1] * d.tx;
            d.wkx = kC3 * wx; d.wky = kC3 * wy; d.wkz = kC3 * wz;
            d.wjx = kCJ * wx; d.wjy = kCJ * wy; d.wjz = kCJ * wz;
            SegF f;
            std::memset(&f, 0, sizeof(f));
            if (L.kind == LOOP_3D) {
                f.bx = (float)d.bx; f.by = (float)d.by; f.bz = (float)d.bz;
                f.tx = (float)d.tx; f.ty = (float)d.ty; f.tz = (float)d.tz;
                f.L = (float)d.L;
                f.nx = (float)d.nx; f.ny = (float)d.ny; f.nz = (float)d.nz;
                f.wkx = (float)d.wkx; f.wky = (float)d.wky; f.wkz = (float)d.wkz;
                f.wjx = (float)d.wjx; f.wjy = (float)d.wjy; f.wjz = (float)d.wjz;
            } else {
                SegP g;
                std::memset(&g, 0, sizeof(g));
                const double e1 = p1[k] - p1[j], e2 = p2[k] - p2[j];
                const double len2 = std::hypot(e1, e2);
                const double t1 = e1 / len2, t2 = e2 / len2;
                g.b1 = (float)p1[k]; g.b2 = (float)p2[k];
                g.t1 = (float)t1; g.t2 = (float)t2;
                g.L = (float)len2;
                g.iL = (float)(1.0 / len2);
                g.e1 = (float)e1;
                g.e2 = (float)e2;
                g.w1 = (float)(-fb3 * t2);
                g.w2 = (float)(fb3 * t1);
                g.w3 = (float)(fb1 * t2 - fb2 * t1);
                g.wt1 = (float)((fb1 * t2 - fb2 * t1) * t1);
                g.wt2 = (float)((fb1 * t2 - fb2 * t1) * t2);
                std::memcpy(&f, &g, sizeof(f));
            }
            if (flat64) {
                SegPD e;
                std::memset(&e, 0, sizeof(e));
                const double e1 = p1[k] - p1[j], e2 = p2[k] - p2[j];
                const double len2 = std::hypot(e1, e2);
                e.b1 = p1[k]; e.b2 = p2[k];
                e.t1 = e1 / len2; e.t2 = e2 / len2;
                e.L = len2;
                e.iL = 1.0 / len2;
                e.e1 = e1;
                e.e2 = e2;
                e.w1 = -b3 * e.t2;
                e.w2 = b3 * e.t1;
                e.w3 = b1 * e.t2 - b2 * e.t1;
                e.wt1 = e.w3 * e.t1;
                e.wt2 = e.w3 * e.t2;
                std::memcpy(&d, &e, sizeof(d));
            }
            out.segd.push_back(d);
            out.segf.push_back(f);
        }
        out.segments += nv;
        out.loops.push_back(L);
    }
    // one far-pass kernel per loop kind (each compiled for its own form)
    std::stable_sort(out.loops.begin(), out.loops.end(),
                     [](const LineLoop& a, const LineLoop& b) { return a.kind < b.kind; });
    for (int k = 0; k <= 4; k++) {
        int c = 0;
        while (c < (int)out.loops.size() && out.loops[c].kind < k) c++;
        out.kofs[k] = c;
    }
    return "";
}

inline std::string pack_loopon(const disco_loopon_loop* in, int nloops, double Lx, double Ly,
                               double Lz, LooponPack& out)
{
    out = LooponPack();
    out.work.assign(nloops, 0.0);
    const double k1 = kC3;
    for (int l = 0; l < nloops; l++) {
        const disco_loopon_loop& lp = in[l];
        if (lp.np < 0) return "negative loopon count";
        if (lp.np > 0 && (!lp.x || !lp.y || !lp.z)) return "null loopon array";
        out.work[l] = (double)lp.np;
        if (lp.np == 0) continue;
        double cx = 0, cy = 0, cz = 0;
        for (int p = 0; p < lp.np; p++) { cx += lp.x[p]; cy += lp.y[p]; cz += lp.z[p]; }
        cx /= lp.np; cy /= lp.np; cz /= lp.np;
        LooponLoop L;
        std::memset(&L, 0, sizeof(L));
        L.cx = cx; L.cy = cy; L.cz = cz;
        const double* ph = lp.phys;
        for (int i = 0; i < 10; i++)
            if (!std::isfinite(ph[i])) return "non-finite loop physics";
        L.bx = ph[0]; L.by = ph[1]; L.bz = ph[2];
        L.mx = ph[3]; L.my = ph[4]; L.mz = ph[5];
        L.nx = ph[6]; L.ny = ph[7]; L.nz = ph[8];
        L.c1 = -k1 * ph[9] * (1.0 - 2.0 * kNu);
        L.c2 = -3.0 * k1 * ph[9];
        L.b_n = ph[3] * ph[6] + ph[4] * ph[7] + ph[5] * ph[8];
        L.p0 = (int)out.pd.size();
        L.np = lp.np;
        double r2max = 0.0;
        for (int p = 0; p < lp.np; p++) {
            const double x = lp.x[p] - cx, y = lp.y[p] - cy, z = lp.z[p] - cz;
            if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z))
                return "non-finite loopon";
            r2max = std::fmax(r2max, x * x + y * y + z * z);
            double4 d; d.x = x; d.y = y; d.z = z; d.w = 0.0;
            float4 f; f.x = (float)x; f.y = (float)y; f.z = (float)z; f.w = 0.0f;
            out.pd.push_back(d);
            out.pf.push_back(f);
        }
        L.rmax2 = (std::sqrt(r2max) + 1e-3) * (std::sqrt(r2max) + 1e-3);

        // flat loop: every loopon within float resolution of the plane through
        // the centroid normal to n; the FP32 table then holds (P.e1, P.e2, 0, 0)
        const double nmag = std::sqrt(L.nx * L.nx + L.ny * L.ny + L.nz * L.nz);
        bool flat = nmag > 0.0;
        if (flat) {
            L.nmag = nmag;
            L.ux = L.nx / nmag; L.uy = L.ny / nmag; L.uz = L.nz / nmag;
            const double a[3] = {std::fabs(L.ux) < 0.9 ? 1.0 : 0.0, std::fabs(L.ux) < 0.9 ? 0.0 : 1.0, 0.0};
            double ex = L.uy * a[2] - L.uz * a[1], ey = L.uz * a[0] - L.ux * a[2],
                   ez = L.ux * a[1] - L.uy * a[0];
            const double en = std::sqrt(ex * ex + ey * ey + ez * ez);
            L.e1x = ex / en; L.e1y = ey / en; L.e1z = ez / en;
            L.e2x = L.uy * L.e1z - L.uz * L.e1y;
            L.e2y = L.uz * L.e1x - L.ux * L.e1z;
            L.e2z = L.ux * L.e1y - L.uy * L.e1x;
            const double res = std::ldexp(std::fmax(std::sqrt(r2max), 1.0), -24);
            const size_t base = out.pd.size() - (size_t)lp.np;
            for (int p = 0; p < lp.np && flat; p++) {
                const double4& d = out.pd[base + p];
                flat = std::fabs(d.x * L.ux + d.y * L.uy + d.z * L.uz) <= res;
            }
            if (flat) {
                for (int p = 0; p < lp.np; p++) {
                    const double4& d = out.pd[base + p];
                    float4& f = out.pf[base + p];
                    f.x = (float)(d.x * L.e1x + d.y * L.e1y + d.z * L.e1z);
                    f.y = (float)(d.x * L.e2x + d.y * L.e2y + d.z * L.e2z);
                    f.z = 0.0f; f.w = 0.0f;
                }
                L.X1 = Lx * L.e1x; L.X2 = Lx * L.e2x; L.X3 = Lx * L.ux;
                L.Y1 = Ly * L.e1y; L.Y2 = Ly * L.e2y; L.Y3 = Ly * L.uy;
                L.Z1 = Lz * L.e1z; L.Z2 = Lz * L.e2z; L.Z3 = Lz * L.uz;
                L.m1 = (float)(L.c2 * (L.mx * L.e1x + L.my * L.e1y + L.mz * L.e1z));
                L.m2 = (float)(L.c2 * (L.mx * L.e2x + L.my * L.e2y + L.mz * L.e2z));
                L.m3 = (float)(L.c2 * (L.mx * L.ux + L.my * L.uy + L.mz * L.uz));
            }
        }
        L.flat = flat ? 1 : 0;
        out.loopons += lp.np;
        out.loops.push_back(L);
        out.src.push_back(l);
    }
    return "";
}

}  // namespace disco
