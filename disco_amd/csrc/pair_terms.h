// pair_terms.h -- per-(atom, loop, image) field evaluations of the hot path.
//
// Two fields, each in FP64 and FP32:
//   * the closed-form line integral of a closed polyline (the Volterra
//     field: solid-angle term b Omega / 4 pi plus the segment terms J0, K0, K1),
//     the quantity of Code/barnett.hpp:44-107;
//   * the Kroupa loopon sum, the quantity of Code/kroupa.hpp:52-73.
//
// The formulas are the reference's, rearranged for a 64-lane VALU:
//   * Omega: the fan of Van Oosterom-Strackee triangles (c, A, B) gives
//     Omega = 2 sum_k atan2(num_k, den_k) = 2 arg prod_k (den_k + i num_k).
//     For a simple flat contour the signed sum lies in (-2 pi, 2 pi) (its
//     |Omega/2| < pi), so the principal argument of the running complex
//     product is exact: one atan2 per loop image instead of one per segment.
//     Any other contour (off its plane, or winding twice round some point of
//     it) can reach |Omega| up to 4 pi: its walk counts the running product's
//     crossings of the negative real axis, each worth 2 pi of the sum (a factor
//     turns the product by less than pi, so the crossing's side and direction
//     follow from the signs of Im before and after and of the turn's cross
//     product -- the reference's per-triangle sum).  num_k = (c - p).(A x B)
//     with A, B relative to c: the cross product A x B is a per-segment
//     constant (det[c-p, A-p, B-p] = det[c-p, A-c, B-c]).  Each factor is
//     scaled by 1/(|c-p||A-p||B-p|) (a positive factor does not move the
//     argument) and the product is renormalised by its binary exponent.
//   * J0 = ln[(R_B + s_B)/(R_A + s_A)] = 2 atanh(L / (R_A + R_B)) for every sign
//     of s_A, s_B (the classical segment identity), evaluated without
//     cancellation: a Taylor series of atanh for small arguments, the log form
//     otherwise.
//   * K0 = [s / (d^2 R)]_A^B: mixed signs of s_A, s_B use the reference's
//     (s_B/R_B - s_A/R_A)/d^2; same signs use the cancellation-free
//     L (s_A + s_B) / (R_A R_B (s_B R_A + s_A R_B)), equal to the reference's
//     two same-sign branches (Code/barnett.hpp:89-92).
//   * d.w = a.w and d K0 + t K1 = a K0 + t (K1 - s_A K0) since w = b x t is
//     normal to t: the normal offset d is needed only for the mixed K0.
// The Kroupa term is regrouped the same way: per image, with g = C1 / rho^3 and
// h = C2 (m.rho)(n.rho) / rho^5, u_i = b_i [n_i S1 + m_i S2 + S3_i],
// S1 = sum g m.rho, S2 = sum g n.rho, S3 = sum rho (h - g m.n) -- which is
// Code/kroupa.hpp:69-72 summed over loopons (C1, C2 folded into constants).
//
// The templates take the math operations as a policy class M (the HIP build
// instantiates them with the gfx950 intrinsics in disco_hip.hip).
#pragma once

#include <cstdint>

#ifndef DISCO_HD
#define DISCO_HD inline
#endif

// DISCO_HOLD("+v"(a), ...) ends the computation of a, ... at this point of
// the device program: an empty asm statement that reads and writes them, so
// the compiler can neither defer their arithmetic past later code nor hoist
// it.  Nothing on the host.
#if defined(__HIP_DEVICE_COMPILE__)
#define DISCO_HOLD(...) asm volatile("" : __VA_ARGS__)
#else
#define DISCO_HOLD(...) ((void)0)
#endif

namespace disco {

constexpr double kPi = 3.14159265358979323846;
constexpr double kNu = 0.35;                                   // Code/kroupa.hpp:24
constexpr double kC3 = 1.0 / (8.0 * kPi * (1.0 - kNu));        // k1, Code/kroupa.hpp:38-41
constexpr double kCJ = kC3 - 1.0 / (4.0 * kPi);                // Code/barnett.hpp:51
constexpr double kInv4Pi = 1.0 / (4.0 * kPi);

// ------------------------------------------------------------ packed tables
// Segment k of a loop runs A = V[k-1] -> B = V[k] (V[nv-1] -> V[0] closes).
// All positions are relative to the loop's apex c.
template <typename R>
struct alignas(16 * sizeof(R)) Seg {
    R bx, by, bz;     // B - c
    R tx, ty, tz;     // unit tangent of A -> B
    R L;              // |B - A|
    R nx, ny, nz;     // (A - c) x (B - c)
    R wkx, wky, wkz;  // cK (b x t)
    R wjx, wjy, wjz;  // cJ (b x t)
};
using SegF = Seg<float>;
using SegD = Seg<double>;

// One loop of a line plan (read with scalar loads: every field is wave-uniform).
struct alignas(64) LineLoop {
    double cx, cy, cz;      // apex
    double bx, by, bz;      // Burgers vector
    double nx, ny, nz;      // unit plane normal
    double e1x, e1y, e1z;   // unit in-plane axes: (e1, e2, n) is the loop frame
    double e2x, e2y, e2z;
    double X1, X2, X3;      // box edges Lx ex, Ly ey, Lz ez in the loop frame
    double Y1, Y2, Y3;
    double Z1, Z2, Z3;
    double rn2;             // near test: in-plane radius^2 bound
    double hn;              // near test: height bound
    float bf[3];            // b, e1, e2, n rounded to float (frame -> global axes
    float e1f[3];           // of the FP32 image fields)
    float e2f[3];
    float nf[3];
    int seg0, nv;           // first segment in the tables, count
    int kind;               // LOOP_3D / LOOP_FLAT / LOOP_GLIDE / LOOP_PRISM
    int flat64;             // flat to double resolution: the FP64 table holds SegPD
    double reach;           // > sqrt(rn2 + hn^2): a near image has its apex this close
    double iLx, iLy, iLz;   // 1 / box lengths (0 for a zero length)
    // 9-point cubature of the flat cut (the far-image form of a cubature plan):
    // weights (areas) of the centre and of the eight ring points at angles
    // j pi/4 from e1 and radius rrho
    float rw0, rw1, rrho;
    float rmp[4];           // rm . P for the ring directions (1, 0), (0, 1), (1, 1)/sqrt2,
                            // (-1, 1)/sqrt2 (times rrho)
    float rm[3];            // c2 m in the frame, c2 = -3 k1 per unit area
    float rcb;              // c1 (m.n), c1 = -k1 (1 - 2 nu) per unit area
    float rmf[3];           // m = b / |b| in the frame
    float rbm;              // |b|
    float rt0, rwinv;       // window: w(r^2) = s(clamp((r^2 - rt0) rwinv, 0, 1)), s quintic
    int rule_ok;            // the cubature matches the cut's moments through degree 5
    float rR;               // the cut's radius: largest in-plane vertex distance from c
    int src;                // index of the loop in the caller's loop array
    // Two-ring cubature of the cut (the band form of a cubature plan): the
    // centre (area bw0) and two rings of twelve points (areas bw1, bw2 each) at
    // radii brho1, brho2, ring 1 at angles j pi/6 from e1, ring 2 at
    // j pi/6 + pi/12; it stands in for the exact field of lanes with
    // r^2 >= rtb (a huge rtb: off)
    float bw0, bw1, bw2;
    float brho1, brho2;
    float bmp1[6], bmp2[6]; // c2 m . P for the six pair directions of each ring
    float rtb;
    int band_ok;            // the two-ring rule matches the cut's moments through degree 9
    int sh0;                // first FP64-table slot of the loop's image shifts (plans)
    // Images (i, j, k) of [-1, 1]^3 left out of the sum (the caller sums them
    // another way): bit 9 (i + 1) + 3 (j + 1) + (k + 1) of oct_skip[o] for the
    // atoms of octant o = [x >= osx] + 2 [y >= osy] + 4 [z >= osz]; skip_all is
    // the AND of the eight masks (images left out for every atom), skip_any
    // their OR (0: nothing left out)
    double osx, osy, osz;
    unsigned oct_skip[8];
    unsigned skip_all, skip_any;
    int rpad2[8];
};
static_assert(sizeof(LineLoop) == 576, "loop record size");

// One loop of a loopon plan; loopons are stored relative to (cx, cy, cz).
struct alignas(128) LooponLoop {
    double cx, cy, cz;   // loopon centroid
    double bx, by, bz;   // phys.b (per-component |b|, Code/kroupa.hpp:30)
    double mx, my, mz;
    double nx, ny, nz;
    double c1, c2;       // -k1 dS (1 - 2 nu), -3 k1 dS
    double b_n;          // m.n
    double rmax2;        // (largest |loopon - centroid| + 1e-3)^2: a lane farther from
                         // the centroid cannot sit on a loopon
    // flat loops (loopons in the plane normal to n, to float resolution): the
    // frame (e1, e2, n/|n|), the box edges in it, |n| and c2 m in it
    double e1x, e1y, e1z, e2x, e2y, e2z, ux, uy, uz;
    double X1, X2, X3, Y1, Y2, Y3, Z1, Z2, Z3;
    double nmag;
    float m1, m2, m3;
    int flat;            // the FP32 loopon table holds frame coordinates (P1, P2, 0)
    int p0, np;
};
static_assert(sizeof(LooponLoop) == 384, "loopon loop record size");

template <typename R>
struct Vec3 { R x, y, z; };

// Segment of a flat loop in the loop's frame (e1, e2, n), e1 and e2 spanning
// its plane: every vertex has zero height, so positions and tangents have two
// components, and (A - c) x (B - c) is normal to the plane.  Same size as
// Seg<float> (the FP32 table holds one layout or the other per loop).
struct alignas(64) SegP {
    float b1, b2;       // B - c in the plane
    float t1, t2;       // unit tangent
    float L;            // |B - A|
    float iL;           // 1 / L
    float w1, w2, w3;   // b x t in the frame
    float e1, e2;       // B - A
    float wt1, wt2;     // w3 t (the glide form's be t = (K1 - sA K0) w3 t)
    float pad[3];
};
static_assert(sizeof(SegP) == sizeof(Seg<float>), "one FP32 segment slot per segment");

// The same in FP64, for a loop flat to double resolution (LineLoop::flat64):
// it then takes the FP64 table's slot of the loop instead of Seg<double>.
struct alignas(128) SegPD {
    double b1, b2;      // B - c in the plane
    double t1, t2;      // unit tangent
    double L;           // |B - A|
    double iL;          // 1 / L
    double w1, w2, w3;  // b x t in the frame
    double e1, e2;      // B - A
    double wt1, wt2;    // w3 t
    double pad[3];
};
static_assert(sizeof(SegPD) == sizeof(Seg<double>), "one FP64 segment slot per segment");

// FP32 layout of a loop's segments.  LOOP_3D: Seg<float> in global axes (a
// loop whose vertices leave its plane by more than float resolution).  The
// flat kinds use SegP: LOOP_FLAT any b, LOOP_GLIDE b in the plane (b x t is
// normal: w1 = w2 = 0), LOOP_PRISM b along n (b x t in the plane: w3 = 0).
enum { LOOP_3D = 0, LOOP_FLAT = 1, LOOP_GLIDE = 2, LOOP_PRISM = 3 };

// ------------------------------------------------------------ atanh pieces
// 2 atanh(z) for 0 <= z < 1.  For z < 1/8: z (2 + y p(y)), y = z^2, p the
// degree-1 minimax fit of (2 atanh(z)/z - 2)/y on [0, 1/64] (relative error
// 2.1e-8 in exact arithmetic, 1.3e-7 evaluated in float, as the degree-2 fit
// on [0, 1/16] is); above it the log form.  Every lane evaluates the series
// and the log form runs only when a lane of the wave needs it: one
// wave-uniform branch instead of two exec-masked halves.
template <class M>
DISCO_HD float two_atanh_vote_f(float z)
{
    const float y = z * z;
    const float p = 0.40630186f * y + 0.66663605f;
    float r = z * (y * p + 2.0f);
    const bool big = !(z < 0.125f);
    if (__builtin_expect(M::any(big), 0)) {
        const float l = M::log_f((1.0f + z) * M::rcp_f(1.0f - z));
        r = big ? l : r;
    }
    return r;
}

// atan2(y, x) in float for finite arguments: t = min/max of |x|, |y| in
// [0, 1], atan(t) = t + t s q(s) with s = t^2 and q of degree 7, the
// relative minimax fit of atan on [0, 1] (1.5e-8 in exact arithmetic, 2.4 ulp
// evaluated in float), then the C library's quadrants and the sign of y
// ((+-0, +-0) gives +-0).  A magnitude |(x, y)| in [1e-37, 1e37] keeps the
// reciprocal of the larger one normal.
template <class M>
DISCO_HD float atan2_own_f(float y, float x)
{
    const float ax = M::fabs_f(x), ay = M::fabs_f(y);
    const float t = M::fmin_f(ax, ay) * M::rcp_f(M::fmax_f(M::fmax_f(ax, ay), 1e-37f));
    const float s = t * t;
    float q = 0.00284985849f;
    q = q * s - 0.0160685126f;
    q = q * s + 0.0426913425f;
    q = q * s - 0.0750428066f;
    q = q * s + 0.106409281f;
    q = q * s - 0.142036438f;
    q = q * s + 0.199926198f;
    q = q * s - 0.333330721f;
    float r = (t * s) * q + t;
    r = ay > ax ? 1.57079637f - r : r;
    r = x < 0.0f ? 3.14159274f - r : r;
    return M::copysign_f(r, y);
}

// ------------------------------------------------------------ FP64 elementary functions
// log and atan2 in double from multiply-adds, one reciprocal and exponent
// manipulation: far fewer live registers than the library routines, which set
// the register count (and so the occupancy) of every FP64 line kernel.
// Accuracy: a few units in the last place (checked against the C library on
// the host, tests/test_math_host.py).

// 2 atanh(s) = log((1+s)/(1-s)) by its odd series, |s| <= 1/4 (terms through
// s^27: the first omitted one is below 1e-18 of the sum)
template <class M>
DISCO_HD double two_atanh_series_d(double s)
{
    const double y = s * s;
    double p = 1.0 / 27;
    p = p * y + 1.0 / 25; p = p * y + 1.0 / 23; p = p * y + 1.0 / 21;
    p = p * y + 1.0 / 19; p = p * y + 1.0 / 17; p = p * y + 1.0 / 15;
    p = p * y + 1.0 / 13; p = p * y + 1.0 / 11; p = p * y + 1.0 / 9;
    p = p * y + 1.0 / 7;  p = p * y + 1.0 / 5;  p = p * y + 1.0 / 3;
    return 2.0 * s + (2.0 * s * y) * p;
}

// log(w) for finite w >= 1: w = m 2^e with m in [1/sqrt2, sqrt2), log m =
// 2 atanh((m - 1)/(m + 1)) (|s| <= 0.172), e ln2 in two parts (the high part
// has 20 trailing zero bits, so e ln2_hi is exact)
template <class M>
DISCO_HD double log_ge1_d(double w)
{
    int e = M::exponent_d(w);                  // w = f 2^e, f in [1/2, 1)
    double m = M::ldexp_d(w, -e);
    if (m < 0.70710678118654752) { m = m * 2.0; e -= 1; }
    const double sr = (m - 1.0) * M::rcp_d(m + 1.0);
    const double ln2_hi = 0.6931471804855391, ln2_lo = 7.440617110012397e-11;
    return e * ln2_hi + (e * ln2_lo + two_atanh_series_d<M>(sr));
}

// atan2(y, x) with the C library's quadrants and signed zeros for finite
// arguments: t = min/max of |x|, |y| in [0, 1], reduced about the nearest
// tan(j pi/16) (j = 0..4) to |u| <= tan(pi/32), then the odd Taylor series of
// atan(u) through u^17 (first omitted term < 5e-18); atan(tan(j pi/16)) as
// stored (each table value tau_j is a double, and its arctangent is carried
// in two parts, computed to 40 digits).
template <class M>
DISCO_HD double atan2_own_d(double y, double x)
{
    const double ax = M::fabs_d(x), ay = M::fabs_d(y);
    const bool swap = ay > ax;
    const double num = swap ? ax : ay, den = swap ? ay : ax;
    const double t = den > 0.0 ? num * M::rcp_d(den) : 0.0;
    double tau = 0.0, ahi = 0.0, alo = 0.0;
    if (t > 0.09849140335716425) { tau = 0.198912367379658; ahi = 0.19634954084936207; alo = 6.846802412842648e-18; }
    if (t > 0.3033466836073424) { tau = 0.41421356237309503; ahi = 0.39269908169872414; alo = 3.060132146563891e-18; }
    if (t > 0.5345111359507916) { tau = 0.6681786379192989; ahi = 0.5890486225480862; alo = -5.412106055318847e-18; }
    if (t > 0.8206787908286602) { tau = 1.0; ahi = 0.7853981633974483; alo = 3.061616997868383e-17; }
    const double u = (t - tau) * M::rcp_d(1.0 + t * tau);
    const double v = u * u;
    double p = -1.0 / 17;
    p = p * v + 1.0 / 15; p = p * v - 1.0 / 13; p = p * v + 1.0 / 11;
    p = p * v - 1.0 / 9;  p = p * v + 1.0 / 7;  p = p * v - 1.0 / 5;
    p = p * v + 1.0 / 3;
    double a = ahi + (alo + (u - (u * v) * p));
    const double pi2_hi = 1.5707963267948966, pi2_lo = 6.123233995736766e-17;
    const double pi_hi = 3.141592653589793, pi_lo = 1.2246467991473532e-16;
    if (swap) a = pi2_hi + (pi2_lo - a);
    if (x < 0.0 || (x == 0.0 && M::signbit_d(x))) a = pi_hi + (pi_lo - a);
    return M::signbit_d(y) ? -a : a;
}

// 2 atanh(z) in double: for z < 1/4 the odd Taylor series through z^27
// (first omitted term y^14/29 < 1e-18 relative, y = z^2 <= 1/16), else
// log1p(2z / (1 - z)).
template <class M>
DISCO_HD double two_atanh_d(double z)
{
    if (z < 0.25) {
        const double y = z * z;
        double p = 1.0 / 27;
        p = p * y + 1.0 / 25; p = p * y + 1.0 / 23; p = p * y + 1.0 / 21;
        p = p * y + 1.0 / 19; p = p * y + 1.0 / 17; p = p * y + 1.0 / 15;
        p = p * y + 1.0 / 13; p = p * y + 1.0 / 11; p = p * y + 1.0 / 9;
        p = p * y + 1.0 / 7;  p = p * y + 1.0 / 5;  p = p * y + 1.0 / 3;
        return 2.0 * z + (2.0 * z * y) * p;
    }
    return log_ge1_d<M>((1.0 + z) * M::rcp_d(1.0 - z));
}

// ------------------------------------------------------------ winding of the product
// +1 when the running product (X, Y) -> (Xn, Yn) crossed the negative real
// axis turning counter-clockwise, -1 clockwise, else 0 (Y >= 0 is the upper
// side, as the principal argument (-pi, pi] takes it).  A turn by less than pi
// from the upper side to the lower crosses the negative axis exactly when it
// turns counter-clockwise, X Yn - Xn Y > 0; near that axis (X, Xn < 0) the two
// products have the same sign, so rounding cannot flip it.
template <typename R>
DISCO_HD int winding_step(R X, R Y, R Xn, R Yn)
{
    const R cr = X * Yn - Xn * Y;
    const bool up = Y >= R(0), upn = Yn >= R(0);
    return (up && !upn && cr > R(0)) ? 1 : ((!up && upn && cr < R(0)) ? -1 : 0);
}

// ------------------------------------------------------------ line, FP32
// Field of one loop image in float; q = atom - c - delta (float), s the loop's
// segments, b the Burgers vector.  Adds into o (float).
template <class M>
DISCO_HD void line_image_f32(float qx, float qy, float qz, const SegF* __restrict__ s, int nv,
                             float bx, float by, float bz, float& ox, float& oy, float& oz)
{
    const float l1sq = qx * qx + qy * qy + qz * qz;
    const float rl1 = M::rsqrt_f(l1sq);
    const float l1 = l1sq * rl1;

    const SegF& last = s[nv - 1];
    float ax = last.bx - qx, ay = last.by - qy, az = last.bz - qz;
    float la2 = ax * ax + ay * ay + az * az;
    float rA = M::rsqrt_f(la2);
    float RA = la2 * rA;
    float P = l1 * RA - (qx * ax + qy * ay + qz * az);        // l1 |A-p| + (c-p).(A-p)

    float X = 1.0f, Y = 0.0f;
    int wind = 0;
    float swx = 0.f, swy = 0.f, swz = 0.f;
    float sax = 0.f, say = 0.f, saz = 0.f;
    float stx = 0.f, sty = 0.f, stz = 0.f;
    // blocks of 60 segments: each scaled factor has modulus <= 4, so the
    // product of 60 stays below 4^60 ~ 1.3e36; renormalised between blocks
    // (atan2 takes the last block's as it is)
    for (int k0 = 0; k0 < nv; k0 += 60) {
        const int k1 = nv < k0 + 60 ? nv : k0 + 60;
        for (int k = k0; k < k1; k++) {
            const SegF& g = s[k];
            const float px = g.bx - qx, py = g.by - qy, pz = g.bz - qz;   // B - p
            const float lb2 = px * px + py * py + pz * pz;
            const float rB = M::rsqrt_f(lb2);
            const float RB = lb2 * rB;
            const float rArB = rA * rB;

            // solid-angle factor den + i num of triangle (c, A, B), scaled by
            // 1 / (|c - p| |A - p| |B - p|)
            const float r1b = -(qx * px + qy * py + qz * pz);              // (c-p).(B-p)
            const float ab = ax * px + ay * py + az * pz;
            const float sc = rArB * rl1;
            const float den = (RB * P + RA * r1b + l1 * ab) * sc;
            const float num = -(qx * g.nx + qy * g.ny + qz * g.nz) * sc;
            P = l1 * RB + r1b;
            const float Xn = X * den - Y * num;
            const float Yn = X * num + Y * den;
            wind += winding_step(X, Y, Xn, Yn);
            X = Xn;
            Y = Yn;

            // segment terms; K0 = kn / kd, one reciprocal shared with J0's argument
            const float sA = ax * g.tx + ay * g.ty + az * g.tz;
            const float sB = sA + g.L;
            const float RS = RA + RB;
            float kn = g.L * (sA + sB) * rArB;
            float kd = sB * RA + sA * RB;
            const bool mixed = (sA < 0.0f) && (sB > 0.0f);      // atom beside the segment
            if (M::any(mixed)) {
                const float dx = ax - sA * g.tx, dy = ay - sA * g.ty, dz = az - sA * g.tz;
                const float d2 = dx * dx + dy * dy + dz * dz;
                const float knm = sB * rB - sA * rA;
                kd = mixed ? d2 : kd;
                kn = mixed ? knm : kn;
            }
            const float rr = M::rcp_f(RS * kd);
            const float J0 = two_atanh_vote_f<M>(g.L * kd * rr);
            const float K0 = kn * RS * rr;
            const float K1 = rA - rB;
            const float dw = ax * g.wkx + ay * g.wky + az * g.wkz;   // cK (d.w) = cK (a.w)
            const float al = dw * K0;
            const float be = dw * (K1 - sA * K0);
            swx += J0 * g.wjx; swy += J0 * g.wjy; swz += J0 * g.wjz;
            sax += al * ax;    say += al * ay;    saz += al * az;
            stx += be * g.tx;  sty += be * g.ty;  stz += be * g.tz;

            ax = px; ay = py; az = pz;
            RA = RB; rA = rB;
        }
        if (k1 < nv) {
            const int e = M::exponent_f(M::fmax_f(M::fabs_f(X), M::fabs_f(Y)));
            X = M::ldexp_f(X, -e);
            Y = M::ldexp_f(Y, -e);
        }
    }
    const float fo = 2.0f * atan2_own_f<M>(Y, X) * (float)kInv4Pi + (float)wind;
    ox += bx * fo + (swx - sax - stx);
    oy += by * fo + (swy - say - sty);
    oz += bz * fo + (swz - saz - stz);
}

// ------------------------------------------------------------ line, FP32, flat loops
// The same algebra in the loop's frame, where the lane sits at q = (q1, q2, h)
// and every vertex at height 0: B - p = (B1 - q1, B2 - q2, -h), so dot
// products with positions and tangents take two products plus a per-image
// constant.  The solid angle of a flat contour is that of its region whatever
// the in-plane apex of the fan, so the walk fans it from the lane's foot
// F = (q1, q2, 0) instead of c: with c - p = (0, 0, -h), l1 = |h|, Van
// Oosterom's factor of triangle (F, A, B) is |h| [RA RB + |h| (RA + RB) + a.b
// - i sgn(h) (a x b).n] (a = A - p, b = B - p), so per segment, scaled by
// 1 / (|h| RA RB), the walk multiplies g = 1 + (|h| RS + a.b) / (RA RB)
// + i (a x b).n / (RA RB) (|g| <= 4) and Omega = -sgn(h) 2 arg prod g, with
// (a x b).n = a x (B - A) from the per-segment edge B - A: free of the
// cancellation of a1 b2 - a2 b1 (terms |q|^2 for a result |q| L) for a lane
// far from the loop.
// A lane at h = 0 outside the contour gets 0, as it should; one on the cut
// itself (h = 0 inside) is in the near slab, which the FP64 pass evaluates.
struct FlatWalk {
    float a1, a2, la2, RA, rA;               // A - p in the plane, |A-p|^2, |A-p|, 1/|A-p|
    float X, Y;                              // running solid-angle product
    float sw1, sw2, sw3;                     // sum J0 w
    float x1, x2, sal;                       // sum (al a + be t) in the plane, sum al (a3 = -h)
};

template <class M, int KIND>
DISCO_HD void flat_step(const SegP& g, float q1, float q2, float h2, float ah, float mh,
                        FlatWalk& w)
{
    const float p1 = g.b1 - q1, p2 = g.b2 - q2;                 // B - p, height -h
    const float lb2 = p1 * p1 + (p2 * p2 + h2);
    const float rB = M::rsqrt_f(lb2);
    const float RB = lb2 * rB;
    const float rArB = w.rA * rB;
    const float RS = w.RA + RB;

    // b = a + L t: a.b = |a|^2 + L sA (|a|^2 is the previous segment's |b|^2,
    // computed from its coordinates, so nothing accumulates along the walk)
    const float sA = w.a1 * g.t1 + w.a2 * g.t2;
    const float den = ((ah * RS + w.la2) + g.L * sA) * rArB + 1.0f;
    const float num = (w.a1 * g.e2 - w.a2 * g.e1) * rArB;
    const float Xn = w.X * den - w.Y * num;
    w.Y = w.X * num + w.Y * den;
    w.X = Xn;

    // K0 = L kn / kd, J0 = 2 atanh(L / RS): one reciprocal of RS kd serves
    // both; kd = sB RA + sA RB = sA RS + L RA
    const float s2 = 2.0f * sA + g.L;
    float kn = s2 * rArB;
    float kd = sA * RS + g.L * w.RA;
    const bool mixed = M::fabs_f(s2) < g.L;                     // sA < 0 < sB
    if (__builtin_expect(M::any(mixed), 0)) {                   // a branch, not selects
        const float sB = sA + g.L;
        const float d1 = w.a1 - sA * g.t1, d2 = w.a2 - sA * g.t2;
        const float dd = d1 * d1 + (d2 * d2 + h2);
        const float knm = (sB * rB - sA * w.rA) * g.iL;
        kd = mixed ? dd : kd;
        kn = mixed ? knm : kn;
    }
    const float Lrr = g.L * M::rcp_f(RS * kd);
    const float J0 = two_atanh_vote_f<M>(kd * Lrr);
    const float K0 = kn * (RS * Lrr);
    const float K1 = w.rA - rB;
    // d.w = a.w with a = (a1, a2, -h); LOOP_GLIDE: a.w = -h w3, the factor -h
    // is applied once after the walk
    float dw;
    if (KIND == LOOP_GLIDE) dw = g.w3;
    else if (KIND == LOOP_PRISM) dw = w.a1 * g.w1 + w.a2 * g.w2;
    else dw = w.a1 * g.w1 + (w.a2 * g.w2 + mh * g.w3);
    const float al = dw * K0;
    if (KIND != LOOP_GLIDE) { w.sw1 += J0 * g.w1; w.sw2 += J0 * g.w2; }
    if (KIND != LOOP_PRISM) w.sw3 += J0 * g.w3;
    if (KIND == LOOP_GLIDE) {
        // dw = w3 is per segment: be t = (K1 - sA K0) (w3 t) from the table
        const float c = K1 - sA * K0;
        w.x1 = (w.x1 + al * w.a1) + c * g.wt1;
        w.x2 = (w.x2 + al * w.a2) + c * g.wt2;
    } else {
        const float be = dw * (K1 - sA * K0);
        w.x1 = (w.x1 + al * w.a1) + be * g.t1;                 // two FMAs each
        w.x2 = (w.x2 + al * w.a2) + be * g.t2;
    }
    w.sal += al;
    // the step's sums end here: deferred past the next step's branches (as
    // the compiler otherwise does) they keep this step's A - p alive beside
    // the next one's, which costs register copies at the loop's back edge
    if (KIND == LOOP_GLIDE) DISCO_HOLD("+v"(w.X), "+v"(w.Y), "+v"(w.x1), "+v"(w.x2), "+v"(w.sw3), "+v"(w.sal));
    else if (KIND == LOOP_PRISM) DISCO_HOLD("+v"(w.X), "+v"(w.Y), "+v"(w.x1), "+v"(w.x2), "+v"(w.sw1), "+v"(w.sw2), "+v"(w.sal));
    else DISCO_HOLD("+v"(w.X), "+v"(w.Y), "+v"(w.x1), "+v"(w.x2), "+v"(w.sw1), "+v"(w.sw2), "+v"(w.sw3), "+v"(w.sal));

    w.a1 = p1; w.a2 = p2;
    w.la2 = lb2; w.RA = RB; w.rA = rB;
}

// Field of one image of a flat loop, in the loop frame: fo = Omega / 4 pi
// (the caller adds b fo) and o = cJ sum J0 w - cK sum (al a + be t).
template <class M, int KIND>
DISCO_HD void line_image_flat_f32(float q1, float q2, float h, const SegP* __restrict__ s, int nv,
                                  float& fo, float& o1, float& o2, float& o3)
{
    const float h2 = h * h, ah = M::fabs_f(h), mh = -h;

    FlatWalk w;
    const SegP& last = s[nv - 1];
    w.a1 = last.b1 - q1;
    w.a2 = last.b2 - q2;
    w.la2 = w.a1 * w.a1 + (w.a2 * w.a2 + h2);
    w.rA = M::rsqrt_f(w.la2);
    w.RA = w.la2 * w.rA;
    w.X = 1.0f; w.Y = 0.0f;
    w.sw1 = w.sw2 = w.sw3 = 0.0f;
    w.x1 = w.x2 = w.sal = 0.0f;
    // blocks of 60 segments (|g| <= 4: a product of 60 stays below 1.3e36),
    // renormalised between blocks; two segments per iteration (fewer
    // register copies of the carried A - p)
    for (int k0 = 0; k0 < nv; k0 += 60) {
        const int k1 = nv < k0 + 60 ? nv : k0 + 60;
        int k = k0;
        for (; k + 1 < k1; k += 2) {
            flat_step<M, KIND>(s[k], q1, q2, h2, ah, mh, w);
            flat_step<M, KIND>(s[k + 1], q1, q2, h2, ah, mh, w);
        }
        for (; k < k1; k++) flat_step<M, KIND>(s[k], q1, q2, h2, ah, mh, w);
        if (k1 < nv) {
            const int e = M::exponent_f(M::fmax_f(M::fabs_f(w.X), M::fabs_f(w.Y)));
            w.X = M::ldexp_f(w.X, -e);
            w.Y = M::ldexp_f(w.Y, -e);
        }
    }
    const float om = 2.0f * atan2_own_f<M>(w.Y, w.X) * (float)kInv4Pi;
    fo = h > 0.0f ? -om : om;
    float x1 = w.x1, x2 = w.x2, x3 = mh * w.sal;
    if (KIND == LOOP_GLIDE) { x1 *= mh; x2 *= mh; x3 *= mh; }
    o1 = (float)kCJ * w.sw1 - (float)kC3 * x1;
    o2 = (float)kCJ * w.sw2 - (float)kC3 * x2;
    o3 = (float)kCJ * w.sw3 - (float)kC3 * x3;
}

// ------------------------------------------------------------ 9-point cubature, FP32
// Far field of a flat loop image from the Kroupa term (Code/kroupa.hpp:52-73)
// integrated over its cut by a 9-point rule: the centre (area rw0) and eight
// ring points P_j (area rw1 each) at radius rho, angles j pi/4, all in the
// loop's plane.  The lane sits at q = (q1, q2, h) in the loop frame; per point
// rho_p = q - P and, with unit-area constants,
//   t1 += rho^-3 (c2 m.rho),  s2 += rho^-3,
//   co  = rho^-3 (h (c2 m.rho) rho^-2 - c1 m.n),  (a1, a2, ac) += co (r1, r2, 1)
// (loopon_flat_f32's sums); the field in the frame is
//   o = |b| [ (0, 0, (c1/c2) t1) + m (c1 h s2) + (a1, a2, h ac) ].
// The ring points come in pairs +-P: |q -+ P|^2 = (|q|^2 + rho^2) -+ 2 q.P and
// c2 m.(q -+ P) = c2 m.q -+ rmp, so a pair shares its dot products, and
// sum co (r1, r2) = (q1, q2) sum co - sum co P: per pair only (co+ - co-) P.
// FIRST: the ring's first pair sets the sums instead of adding to them.
template <class M, bool FIRST = false>
DISCO_HD void rule_ring_pair(float base, float d, float mq, float mp, float h, float cb,
                             float& t1, float& s2, float& ac, float& dc)
{
    const float rp = base - d, rn = base + d;          // |q - P|^2, |q + P|^2
    const float ip = M::rsqrt_f(rp), in = M::rsqrt_f(rn);
    const float ip2 = ip * ip, in2 = in * in;
    const float ip3 = ip2 * ip, in3 = in2 * in;
    const float mrp = mq - mp, mrn = mq + mp;
    const float dt = ip3 * mrp + in3 * mrn, ds = ip3 + in3;
    const float cp = ip3 * ((h * mrp) * ip2 - cb);
    const float cn = in3 * ((h * mrn) * in2 - cb);
    t1 = FIRST ? dt : t1 + dt;
    s2 = FIRST ? ds : s2 + ds;
    ac = FIRST ? cp + cn : ac + (cp + cn);
    dc = cp - cn;
}

template <class M>
DISCO_HD void rule9_image_f32(float q1, float q2, float h, float t, const LineLoop& L,
                              float& o1, float& o2, float& o3)
{
    constexpr float kC1 = (float)(-kC3 * (1.0 - 2.0 * kNu));
    constexpr float kR12 = (float)((1.0 - 2.0 * kNu) / 3.0);     // c1 / c2
    constexpr float kS = 0.70710678f;                            // 1 / sqrt 2
    const float mq = L.rm[0] * q1 + (L.rm[1] * q2 + L.rm[2] * h);
    const float cb = L.rcb;
    // centre: rho = q, |q|^2 = t
    const float i0 = M::rsqrt_f(t);
    const float i02 = i0 * i0, i03 = i02 * i0;
    const float c0 = i03 * ((h * mq) * i02 - cb);
    // ring
    const float rho = L.rrho;
    const float base = t + rho * rho;
    const float u1 = (2.0f * rho) * q1, u2 = (2.0f * rho) * q2;   // 2 q.P on the axes
    const float v1 = (u1 + u2) * kS, v2 = (u2 - u1) * kS;          // and on the diagonals
    float t1, s2, ac, d0, d1, d2, d3;
    rule_ring_pair<M, true>(base, u1, mq, L.rmp[0], h, cb, t1, s2, ac, d0);
    rule_ring_pair<M>(base, u2, mq, L.rmp[1], h, cb, t1, s2, ac, d1);
    rule_ring_pair<M>(base, v1, mq, L.rmp[2], h, cb, t1, s2, ac, d2);
    rule_ring_pair<M>(base, v2, mq, L.rmp[3], h, cb, t1, s2, ac, d3);
    // sum over the ring of co P: pairs along (1,0), (0,1), (1,1)/sqrt2, (-1,1)/sqrt2
    const float A1 = rho * (d0 + (d2 - d3) * kS);
    const float A2 = rho * (d1 + (d2 + d3) * kS);
    const float w0 = L.rw0, w1 = L.rw1;
    const float T1 = w0 * (i03 * mq) + w1 * t1;
    const float S2 = kC1 * (h * (w0 * i03 + w1 * s2));
    const float AC = w0 * c0 + w1 * ac;
    o1 = L.rbm * (L.rmf[0] * S2 + (q1 * AC - w1 * A1));
    o2 = L.rbm * (L.rmf[1] * S2 + (q2 * AC - w1 * A2));
    o3 = L.rbm * ((kR12 * T1 + L.rmf[2] * S2) + h * AC);
}

// The two-ring cubature (LineLoop::bw0 ... bmp2) of the same Kroupa term: the
// centre and rings of twelve at radii brho1 (angles j pi/6) and brho2
// (j pi/6 + pi/12), each ring as six pairs +-P sharing their dot products, as
// in rule9_image_f32.
// cos and sin of the pair directions: ring 0 at j pi/6, ring 1 at j pi/6 + pi/12
constexpr float kBandCos[2][6] = {
    {1.0f, 0.866025403784438647f, 0.5f, 0.0f, -0.5f, -0.866025403784438647f},
    {0.965925826289068287f, 0.707106781186547524f, 0.258819045102520762f,
     -0.258819045102520762f, -0.707106781186547524f, -0.965925826289068287f}};
constexpr float kBandSin[2][6] = {
    {0.0f, 0.5f, 0.866025403784438647f, 1.0f, 0.866025403784438647f, 0.5f},
    {0.258819045102520762f, 0.707106781186547524f, 0.965925826289068287f,
     0.965925826289068287f, 0.707106781186547524f, 0.258819045102520762f}};

template <class M, int RING>
DISCO_HD void band_ring_f32(float q1, float q2, float t, float rho, const float* mp, float mq,
                            float h, float cb, float& t1, float& s2, float& ac, float& A1,
                            float& A2)
{
    const float base = t + rho * rho;
    const float Q1 = (2.0f * rho) * q1, Q2 = (2.0f * rho) * q2;
    float a1, a2, d0;
    if (RING == 0) {                   // first direction (1, 0)
        rule_ring_pair<M, true>(base, Q1, mq, mp[0], h, cb, t1, s2, ac, d0);
        a1 = d0;
        a2 = 0.0f;
    } else {
        const float c = kBandCos[1][0], sn = kBandSin[1][0];
        rule_ring_pair<M, true>(base, c * Q1 + sn * Q2, mq, mp[0], h, cb, t1, s2, ac, d0);
        a1 = d0 * c;
        a2 = d0 * sn;
    }
#pragma unroll
    for (int j = 1; j < 6; j++) {
        const float c = kBandCos[RING][j], sn = kBandSin[RING][j];
        float dc;
        rule_ring_pair<M>(base, c * Q1 + sn * Q2, mq, mp[j], h, cb, t1, s2, ac, dc);
        a1 += dc * c;
        a2 += dc * sn;
    }
    A1 = rho * a1;
    A2 = rho * a2;
}

template <class M>
DISCO_HD void band_image_f32(float q1, float q2, float h, float t, const LineLoop& L,
                             float& o1, float& o2, float& o3)
{
    constexpr float kC1 = (float)(-kC3 * (1.0 - 2.0 * kNu));
    constexpr float kR12 = (float)((1.0 - 2.0 * kNu) / 3.0);     // c1 / c2
    const float mq = L.rm[0] * q1 + (L.rm[1] * q2 + L.rm[2] * h);
    const float cb = L.rcb;
    const float i0 = M::rsqrt_f(t);
    const float i02 = i0 * i0, i03 = i02 * i0;
    const float c0 = i03 * ((h * mq) * i02 - cb);
    float t1a, s2a, aca, A1a, A2a;
    float t1b, s2b, acb, A1b, A2b;
    band_ring_f32<M, 0>(q1, q2, t, L.brho1, L.bmp1, mq, h, cb, t1a, s2a, aca, A1a, A2a);
    band_ring_f32<M, 1>(q1, q2, t, L.brho2, L.bmp2, mq, h, cb, t1b, s2b, acb, A1b, A2b);
    const float w0 = L.bw0, w1 = L.bw1, w2 = L.bw2;
    const float T1 = w0 * (i03 * mq) + (w1 * t1a + w2 * t1b);
    const float S2 = kC1 * (h * (w0 * i03 + (w1 * s2a + w2 * s2b)));
    const float AC = w0 * c0 + (w1 * aca + w2 * acb);
    const float A1 = w1 * A1a + w2 * A1b, A2 = w1 * A2a + w2 * A2b;
    o1 = L.rbm * (L.rmf[0] * S2 + (q1 * AC - A1));
    o2 = L.rbm * (L.rmf[1] * S2 + (q2 * AC - A2));
    o3 = L.rbm * ((kR12 * T1 + L.rmf[2] * S2) + h * AC);
}

// Window weight of the cubature for a lane at squared distance t from the
// loop centre: 0 inside rt0, 1 beyond rt0 + 1/rwinv, the quintic smoothstep
// x^3 (10 - 15 x + 6 x^2) between (a hard switch when rwinv is huge).
DISCO_HD float rule_window(float t, const LineLoop& L)
{
    float x = (t - L.rt0) * L.rwinv;
    x = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
    return (x * x) * x * (x * (x * 6.0f - 15.0f) + 10.0f);
}

// ------------------------------------------------------------ line, FP64
template <class M>
DISCO_HD void line_image_f64(double qx, double qy, double qz, const SegD* __restrict__ s, int nv,
                             double bx, double by, double bz, double& ox, double& oy, double& oz)
{
    const double l1sq = qx * qx + qy * qy + qz * qz;
    const double rl1 = M::rsqrt_d1(l1sq);
    const double l1 = l1sq * rl1;

    const SegD& last = s[nv - 1];
    double ax = last.bx - qx, ay = last.by - qy, az = last.bz - qz;
    const double la2 = ax * ax + ay * ay + az * az;
    double rA = M::rsqrt_d1(la2);
    double RA = la2 * rA;
    double P = l1 * RA + (l1sq - (qx * last.bx + qy * last.by + qz * last.bz));

    double X = 1.0, Y = 0.0;
    int wind = 0;
    double swx = 0., swy = 0., swz = 0., sax = 0., say = 0., saz = 0., stx = 0., sty = 0., stz = 0.;
    for (int k = 0; k < nv; k++) {
        const SegD& g = s[k];
        const double px = g.bx - qx, py = g.by - qy, pz = g.bz - qz;
        const double lb2 = px * px + py * py + pz * pz;
        const double rB = M::rsqrt_d1(lb2);
        const double RB = lb2 * rB;

        const double r1b = l1sq - (qx * g.bx + qy * g.by + qz * g.bz);
        const double ab = ax * px + ay * py + az * pz;
        double den = RB * P + RA * r1b + l1 * ab;
        double num = -(qx * g.nx + qy * g.ny + qz * g.nz);
        P = l1 * RB + r1b;
        const double sc = rl1 * rA * rB;
        den *= sc;
        num *= sc;
        const double Xn = X * den - Y * num;
        const double Yn = X * num + Y * den;
        wind += winding_step(X, Y, Xn, Yn);
        X = Xn;
        Y = Yn;
        if ((k & 63) == 63) {
            const int e = M::exponent_d(M::fmax_d(M::fabs_d(X), M::fabs_d(Y)));
            X = M::ldexp_d(X, -e);
            Y = M::ldexp_d(Y, -e);
        }

        const double sA = ax * g.tx + ay * g.ty + az * g.tz;
        const double sB = sA + g.L;
        const double RS = RA + RB;
        double kn = g.L * (sA + sB) * (rA * rB);
        double kd = sB * RA + sA * RB;
        if ((sA < 0.0) && (sB > 0.0)) {
            const double dx = ax - sA * g.tx, dy = ay - sA * g.ty, dz = az - sA * g.tz;
            kd = dx * dx + dy * dy + dz * dz;
            kn = sB * rB - sA * rA;
        }
        const double rr = M::rcp_d1(RS * kd);
        const double J0 = two_atanh_d<M>(g.L * kd * rr);
        const double K0 = kn * RS * rr;
        const double K1 = rA - rB;
        const double dw = ax * g.wkx + ay * g.wky + az * g.wkz;
        const double al = dw * K0;
        const double be = dw * (K1 - sA * K0);
        swx += J0 * g.wjx; swy += J0 * g.wjy; swz += J0 * g.wjz;
        sax += al * ax;    say += al * ay;    saz += al * az;
        stx += be * g.tx;  sty += be * g.ty;  stz += be * g.tz;

        ax = px; ay = py; az = pz;
        RA = RB; rA = rB;
    }
    const double fo = 2.0 * atan2_own_d<M>(Y, X) * kInv4Pi + (double)wind;
    ox += bx * fo + (swx - sax - stx);
    oy += by * fo + (swy - say - sty);
    oz += bz * fo + (swz - saz - stz);
}

// ------------------------------------------------------------ line, FP64, loop frame
// line_image_flat_f32's walk in FP64 for a loop flat to double resolution:
// the atom at (q1, q2, h) in the loop frame, segments as SegPD, the fan from
// the lane's foot; the general form (any b).  Reciprocal square roots and reciprocals from the hardware
// estimates refined by Newton steps; 2 atanh from its series unless a lane of
// the wave needs the log form.
struct FlatWalkD {
    double a1, a2, la2, RA, rA;
    double X, Y;
    double sw1, sw2, sw3;
    double x1, x2, sal;
};

template <class M>
DISCO_HD double two_atanh_vote_d(double z)
{
    const double y = z * z;
    // z < 1/16: the series to z^13 (the first omitted term, 2 z^15 / 15, is
    // below 2^-59 of the sum; its coefficients are the long series' own, which
    // keeps the constants of the two in the same registers: an economised
    // degree-4 form with one multiply-add fewer measured 4.6 % slower); the one
    // to z^27 for z < 1/4 only when a lane of the wave needs it, so each lane's
    // result depends on its own z alone
    double p = 1.0 / 13;
    p = p * y + 1.0 / 11; p = p * y + 1.0 / 9;
    p = p * y + 1.0 / 7;  p = p * y + 1.0 / 5;  p = p * y + 1.0 / 3;
    const bool mid = !(z < 0.0625);
    if (__builtin_expect(!M::any(mid), 1)) return 2.0 * z + (2.0 * z * y) * p;
    double q = 1.0 / 27;
    q = q * y + 1.0 / 25; q = q * y + 1.0 / 23; q = q * y + 1.0 / 21;
    q = q * y + 1.0 / 19; q = q * y + 1.0 / 17; q = q * y + 1.0 / 15;
    q = q * y + 1.0 / 13; q = q * y + 1.0 / 11; q = q * y + 1.0 / 9;
    q = q * y + 1.0 / 7;  q = q * y + 1.0 / 5;  q = q * y + 1.0 / 3;
    p = mid ? q : p;
    double r = 2.0 * z + (2.0 * z * y) * p;
    // z >= 1/4 implies z >= 1/16: tested only in a wave that has a mid lane
    const bool big = !(z < 0.25);
    if (__builtin_expect(M::any(big), 0)) {                    // a branch, not selects
        const double l = log_ge1_d<M>((1.0 + z) * M::rcp_d(1.0 - z));
        r = big ? l : r;
    }
    return r;
}

// GLIDE: b in the loop plane, so b x t is along n (w1 = w2 = 0) and a.w =
// -h w3: the factor -h is applied once after the walk, as in the FP32 form.
template <class M, bool GLIDE>
DISCO_HD void flat_step_d(const SegPD& g, double q1, double q2, double h2, double ah,
                          double mh, FlatWalkD& w)
{
    const double p1 = g.b1 - q1, p2 = g.b2 - q2;
    const double lb2 = p1 * p1 + (p2 * p2 + h2);
    const double rB = M::rsqrt_d1(lb2);
    const double RB = lb2 * rB;
    const double rArB = w.rA * rB;
    const double RS = w.RA + RB;

    const double sA = w.a1 * g.t1 + w.a2 * g.t2;
    const double den = ((ah * RS + w.la2) + g.L * sA) * rArB + 1.0;
    const double num = (w.a1 * g.e2 - w.a2 * g.e1) * rArB;
    const double Xn = w.X * den - w.Y * num;
    w.Y = w.X * num + w.Y * den;
    w.X = Xn;

    const double s2 = 2.0 * sA + g.L;
    double kn = s2 * rArB;
    double kd = sA * RS + g.L * w.RA;
    const bool mixed = M::fabs_d(s2) < g.L;                     // sA < 0 < sB
    if (__builtin_expect(M::any(mixed), 0)) {
        const double sB = sA + g.L;
        const double d1 = w.a1 - sA * g.t1, d2 = w.a2 - sA * g.t2;
        const double dd = d1 * d1 + (d2 * d2 + h2);
        const double knm = (sB * rB - sA * w.rA) * g.iL;
        kd = mixed ? dd : kd;
        kn = mixed ? knm : kn;
    }
    const double Lrr = g.L * M::rcp_d1(RS * kd);
    const double J0 = two_atanh_vote_d<M>(kd * Lrr);
    const double K0 = kn * (RS * Lrr);
    const double K1 = w.rA - rB;
    const double dw = GLIDE ? g.w3 : w.a1 * g.w1 + (w.a2 * g.w2 + mh * g.w3);
    const double al = dw * K0;
    if (!GLIDE) {
        w.sw1 += J0 * g.w1;
        w.sw2 += J0 * g.w2;
    }
    w.sw3 += J0 * g.w3;
    if (GLIDE) {
        const double c = K1 - sA * K0;
        w.x1 = (w.x1 + al * w.a1) + c * g.wt1;
        w.x2 = (w.x2 + al * w.a2) + c * g.wt2;
    } else {
        const double be = dw * (K1 - sA * K0);
        w.x1 = (w.x1 + al * w.a1) + be * g.t1;
        w.x2 = (w.x2 + al * w.a2) + be * g.t2;
    }
    w.sal += al;
    // the step's sums end here (as in flat_step)
    if (GLIDE) DISCO_HOLD("+v"(w.X), "+v"(w.Y), "+v"(w.x1), "+v"(w.x2), "+v"(w.sw3), "+v"(w.sal));
    else DISCO_HOLD("+v"(w.X), "+v"(w.Y), "+v"(w.x1), "+v"(w.x2), "+v"(w.sw1), "+v"(w.sw2), "+v"(w.sw3), "+v"(w.sal));

    w.a1 = p1; w.a2 = p2;
    w.la2 = lb2; w.RA = RB; w.rA = rB;
}

// Field of one image in the loop frame: fo = Omega / 4 pi (the caller adds
// b fo) and o = cJ sum J0 w - cK sum (al a + be t).
template <class M, bool GLIDE = false>
DISCO_HD void line_image_flat_f64(double q1, double q2, double h, const SegPD* __restrict__ s,
                                  int nv, double& fo, double& o1, double& o2, double& o3)
{
    const double h2 = h * h, ah = M::fabs_d(h), mh = -h;

    FlatWalkD w;
    const SegPD& last = s[nv - 1];
    w.a1 = last.b1 - q1;
    w.a2 = last.b2 - q2;
    w.la2 = w.a1 * w.a1 + (w.a2 * w.a2 + h2);
    w.rA = M::rsqrt_d1(w.la2);
    w.RA = w.la2 * w.rA;
    w.X = 1.0; w.Y = 0.0;
    w.sw1 = w.sw2 = w.sw3 = 0.0;
    w.x1 = w.x2 = w.sal = 0.0;
    for (int k0 = 0; k0 < nv; k0 += 64) {
        const int k1 = nv < k0 + 64 ? nv : k0 + 64;
        // two segments per iteration: no register copies of the carried A - p
        int k = k0;
#pragma unroll 1
        for (; k + 1 < k1; k += 2) {
            flat_step_d<M, GLIDE>(s[k], q1, q2, h2, ah, mh, w);
            flat_step_d<M, GLIDE>(s[k + 1], q1, q2, h2, ah, mh, w);
        }
        if (k < k1) flat_step_d<M, GLIDE>(s[k], q1, q2, h2, ah, mh, w);
        const int e = M::exponent_d(M::fmax_d(M::fabs_d(w.X), M::fabs_d(w.Y)));
        w.X = M::ldexp_d(w.X, -e);
        w.Y = M::ldexp_d(w.Y, -e);
    }
    const double om = 2.0 * atan2_own_d<M>(w.Y, w.X) * kInv4Pi;
    fo = h > 0.0 ? -om : om;
    double x1 = w.x1, x2 = w.x2, x3 = mh * w.sal;
    if (GLIDE) { x1 *= mh; x2 *= mh; x3 *= mh; }
    o1 = kCJ * w.sw1 - kC3 * x1;
    o2 = kCJ * w.sw2 - kC3 * x2;
    o3 = kCJ * w.sw3 - kC3 * x3;
}

// ------------------------------------------------------------ loopons
// Partial sums of one loop image over loopons [0, np), rho = q - P_p:
//   t1 = sum rho^-3 (m'.rho),  t2 = sum rho^-3 (n.rho),
//   s3 = sum rho (m'.rho (n.rho) rho^-5 - cb rho^-3)
// with m' = c2 m and cb = c1 (m.n), so that for the loop
//   S1 = c1 t1 / c2, S2 = c1 t2, S3 = s3 (the sums of the header).
// A pair with rho < 1e-10 adds nothing (Code/kroupa.hpp:57-67): CHECKED
// evaluates that guard; the caller may pass CHECKED = false when no lane of
// the wave lies within the loop's loopon radius (then no pair can be that close).
template <class M, bool CHECKED>
DISCO_HD void loopon_image_f32(float qx, float qy, float qz, const float4* __restrict__ P, int np,
                               float mx, float my, float mz, float nx, float ny, float nz,
                               float cb, float& t1, float& t2, float& s3x, float& s3y,
                               float& s3z)
{
#pragma unroll 4
    for (int p = 0; p < np; p++) {
        const float4 v = P[p];
        const float rx = qx - v.x, ry = qy - v.y, rz = qz - v.z;
        float r2 = rx * rx + ry * ry + rz * rz;
        const bool keep = !CHECKED || r2 >= 1e-20f;
        if (CHECKED) r2 = keep ? r2 : 1.0f;
        const float ri = M::rsqrt_f(r2);
        const float ri2 = ri * ri;
        float ri3 = ri2 * ri;
        const float mr = mx * rx + my * ry + mz * rz;
        const float nr = nx * rx + ny * ry + nz * rz;
        float h = (mr * nr) * (ri3 * ri2);
        if (CHECKED) {
            ri3 = keep ? ri3 : 0.0f;
            h = keep ? h : 0.0f;
        }
        t1 += ri3 * mr;
        t2 += ri3 * nr;
        const float co = h - cb * ri3;
        s3x += co * rx; s3y += co * ry; s3z += co * rz;
    }
}

// The same sums for a flat loop in its frame, the lane at (q1, q2, h): with
// rho = (r1, r2, h), n.rho = |n| h for every loopon, so per loopon
//   t1 += rho^-3 (m'.rho),  s2 += rho^-3,  co = rho^-3 (|n| h (m'.rho) rho^-2 - cb),
//   (s31, s32, sco) += co (r1, r2, 1)
// and the caller forms t2 = |n| h s2 and s3 = (s31, s32, h sco).  m3h = m'_3 h,
// hn = |n| h, h2 = h^2.
template <class M, bool CHECKED>
DISCO_HD void loopon_flat_f32(float q1, float q2, float h2, const float4* __restrict__ P, int np,
                              float m1, float m2, float m3h, float hn, float cb, float& t1,
                              float& s2, float& s31, float& s32, float& sco)
{
#pragma unroll 4
    for (int p = 0; p < np; p++) {
        const float4 v = P[p];
        const float r1 = q1 - v.x, r2 = q2 - v.y;
        float rr = r1 * r1 + (r2 * r2 + h2);
        const bool keep = !CHECKED || rr >= 1e-20f;
        if (CHECKED) rr = keep ? rr : 1.0f;
        const float ri = M::rsqrt_f(rr);
        const float ri2 = ri * ri;
        float ri3 = ri2 * ri;
        if (CHECKED) ri3 = keep ? ri3 : 0.0f;
        const float mr = m1 * r1 + (m2 * r2 + m3h);
        t1 += ri3 * mr;
        s2 += ri3;
        const float co = ri3 * (hn * (mr * ri2) - cb);
        s31 += co * r1; s32 += co * r2; sco += co;
    }
}

template <class M>
DISCO_HD void loopon_image_f64(double qx, double qy, double qz, const double4* __restrict__ P,
                               int np, double mx, double my, double mz, double nx, double ny,
                               double nz, double b_n, double c1, double c2, double& s1,
                               double& s2, double& s3x, double& s3y, double& s3z)
{
    for (int p = 0; p < np; p++) {
        const double4 v = P[p];
        const double rx = qx - v.x, ry = qy - v.y, rz = qz - v.z;
        const double r2 = rx * rx + ry * ry + rz * rz;
        const bool keep = !(r2 < 1e-20);                // r < 1e-10, Code/kroupa.hpp:59
        const double ri = M::rsqrt_d1(keep ? r2 : 1.0);
        const double ri2 = ri * ri;
        const double ri3 = ri2 * ri;
        const double mr = mx * rx + my * ry + mz * rz;
        const double nr = nx * rx + ny * ry + nz * rz;
        const double g = keep ? c1 * ri3 : 0.0;
        const double h = keep ? c2 * (mr * nr) * (ri3 * ri2) : 0.0;
        s1 += g * mr;
        s2 += g * nr;
        const double co = h - g * b_n;
        s3x += co * rx; s3y += co * ry; s3z += co * rz;
    }
}

}  // namespace disco
