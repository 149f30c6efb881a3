// pair_terms.h -- per-(atom, loop, image) field evaluations of the hot path.
//
// Two fields, each in FP64 and FP32:
//   * the closed-form line integral of a closed planar polyline (the Volterra
//     field: solid-angle term b Omega / 4 pi plus the segment terms J0, K0, K1),
//     the quantity of Code/barnett.hpp:44-107;
//   * the Kroupa loopon sum, the quantity of Code/kroupa.hpp:52-73.
//
// The formulas are the reference's, rearranged for a 64-lane VALU:
//   * Omega: the fan of Van Oosterom-Strackee triangles (c, A, B) gives
//     Omega = 2 sum_k atan2(num_k, den_k) = 2 arg prod_k (den_k + i num_k).
//     For a flat contour the signed sum lies in (-2 pi, 2 pi) (its |Omega/2| < pi),
//     so the principal argument of the running complex product is exact: one
//     atan2 per loop image instead of one per segment.  num_k = (c - p).(A x B)
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

// One loop of a line plan.
struct alignas(128) LineLoop {
    double cx, cy, cz;   // apex
    double bx, by, bz;   // Burgers vector
    double nx, ny, nz;   // unit plane normal (near test)
    double rn2;          // near test: in-plane radius^2 bound
    double hn;           // near test: height bound
    int seg0, nv;        // first segment in the tables, count
    double pad[4];
};

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
    int p0, np;
};

template <typename R>
struct Vec3 { R x, y, z; };

// ------------------------------------------------------------ atanh pieces
// 2 atanh(z) for 0 <= z < 1.  Series z (1 + z^2/3 + ... + z^10/11) for
// z < 1/4 (first omitted term z^12/13 < 4.6e-9 relative); log form above.
template <class M>
DISCO_HD float two_atanh_f(float z)
{
    if (z < 0.25f) {
        const float y = z * z;
        float p = 1.0f / 11.0f;
        p = p * y + 1.0f / 9.0f;
        p = p * y + 1.0f / 7.0f;
        p = p * y + 1.0f / 5.0f;
        p = p * y + 1.0f / 3.0f;
        return 2.0f * z + (2.0f * z * y) * p;
    }
    return M::log_f((1.0f + z) * M::rcp_f(1.0f - z));
}

template <class M>
DISCO_HD double two_atanh_d(double z)
{
    return M::log1p_d(2.0 * z / (1.0 - z));
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
    float swx = 0.f, swy = 0.f, swz = 0.f;
    float sax = 0.f, say = 0.f, saz = 0.f;
    float stx = 0.f, sty = 0.f, stz = 0.f;
    // blocks of 16 segments: the complex product is renormalised between
    // blocks (each scaled factor has modulus <= 4, so 16 of them stay finite)
    for (int k0 = 0; k0 < nv; k0 += 16) {
        const int k1 = nv < k0 + 16 ? nv : k0 + 16;
#pragma unroll 2
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
            Y = X * num + Y * den;
            X = Xn;

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
            const float J0 = two_atanh_f<M>(g.L * kd * rr);
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
        const int e = M::exponent_f(M::fmax_f(M::fabs_f(X), M::fabs_f(Y)));
        X = M::ldexp_f(X, -e);
        Y = M::ldexp_f(Y, -e);
    }
    const float fo = 2.0f * M::atan2_f(Y, X) * (float)kInv4Pi;
    ox += bx * fo + (swx - sax - stx);
    oy += by * fo + (swy - say - sty);
    oz += bz * fo + (swz - saz - stz);
}

// ------------------------------------------------------------ line, FP64
template <class M>
DISCO_HD void line_image_f64(double qx, double qy, double qz, const SegD* __restrict__ s, int nv,
                             double bx, double by, double bz, double& ox, double& oy, double& oz)
{
    const double l1sq = qx * qx + qy * qy + qz * qz;
    const double l1 = M::sqrt_d(l1sq);
    const double rl1 = 1.0 / l1;

    const SegD& last = s[nv - 1];
    double ax = last.bx - qx, ay = last.by - qy, az = last.bz - qz;
    double RA = M::sqrt_d(ax * ax + ay * ay + az * az);
    double rA = 1.0 / RA;
    double P = l1 * RA + (l1sq - (qx * last.bx + qy * last.by + qz * last.bz));

    double X = 1.0, Y = 0.0;
    double swx = 0., swy = 0., swz = 0., sax = 0., say = 0., saz = 0., stx = 0., sty = 0., stz = 0.;
    for (int k = 0; k < nv; k++) {
        const SegD& g = s[k];
        const double px = g.bx - qx, py = g.by - qy, pz = g.bz - qz;
        const double RB = M::sqrt_d(px * px + py * py + pz * pz);
        const double rB = 1.0 / RB;

        const double r1b = l1sq - (qx * g.bx + qy * g.by + qz * g.bz);
        const double ab = ax * px + ay * py + az * pz;
        double den = RB * P + RA * r1b + l1 * ab;
        double num = -(qx * g.nx + qy * g.ny + qz * g.nz);
        P = l1 * RB + r1b;
        const double sc = rl1 * rA * rB;
        den *= sc;
        num *= sc;
        const double Xn = X * den - Y * num;
        Y = X * num + Y * den;
        X = Xn;
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
        const double rr = 1.0 / (RS * kd);
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
    const double fo = 2.0 * M::atan2_d(Y, X) * kInv4Pi;
    ox += bx * fo + (swx - sax - stx);
    oy += by * fo + (swy - say - sty);
    oz += bz * fo + (swz - saz - stz);
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
        const double r = M::sqrt_d(r2);
        const bool keep = !(r < 1e-10);                 // Code/kroupa.hpp:59
        const double ri = keep ? 1.0 / r : 1.0;
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
