// geometry.cpp -- host producer of loop inputs (include/disco_geometry.h).
//
// Follows the reference's loop code formula by formula and in its operation
// order: regularPolygon (Code/loop.cpp:585-608), initCommon / defaultAxes
// (:13-57), the loop constructors (:62-244), snapToMidPlane (:342-374) with
// latticePlaneSpacing (Code/lattice.hpp:146-153), loop::polyline and
// roundedPolygon2D (:478-580), generateLoopons / insideUV (:253-316) and
// generateSurface (:426-435), and the Burgers-vector helpers of the prismatic
// slabs: isLatticeTranslation, classifyBurgers and frankCopyShift
// (Code/lattice.hpp:159-168,189-212,239-256).
//
// Rounding: the reference's host build (g++ -O3 -march=native) contracts
// products into fused multiply-adds, so the same formula rounds differently
// with and without FMA hardware.  This file has contraction off and writes
// every fused operation the reference's build performs as an explicit
// std::fma, by GCC's rule for a chain of sums of products: the first product
// of a chain is fused with what it is added to, later products with the
// running sum --
//   a*b + c*d        -> fma(a, b, c*d)
//   a*b + c*d + e*f  -> fma(e, f, fma(a, b, c*d))
//   a*b - c*d        -> fma(a, b, -(c*d))
//   x - a*b          -> fma(-a, b, x),   x + a*b -> fma(a, b, x)
// The same build evaluates sin(x) and cos(x) of one argument as one sincos(x)
// call, whose last bit can differ from sin's or cos's (glibc); those pairs call
// sincos here too.  So the results are bit-identical to the compiled
// reference's (tested against the reference's own sources compiled with its
// flags).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "disco_geometry.h"

#pragma clang fp contract(off)

static const double kPiGeo = 3.14159265358979323846;   // the double nearest pi

namespace {

struct V3 {
    double x, y, z;
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

V3 v3(const double* p) { return {p[0], p[1], p[2]}; }
double dot3(const V3& a, const V3& b) { return std::fma(a.z, b.z, std::fma(a.x, b.x, a.y * b.y)); }
V3 cross(const V3& a, const V3& b)
{
    return {std::fma(a.y, b.z, -(a.z * b.y)), std::fma(a.z, b.x, -(a.x * b.z)),
            std::fma(a.x, b.y, -(a.y * b.x))};
}
V3 scale_inv(const V3& a, double l) { return {a.x / l, a.y / l, a.z / l}; }
double mx(double a, double b) { return a < b ? b : a; }   // std::max
double mn(double a, double b) { return b < a ? b : a; }   // std::min

// in-plane axes (u, v) for unit m, n (Code/loop.cpp:40-57)
void default_axes(const V3& m, const V3& n, V3& u, V3& v)
{
    const double mn_ = dot3(m, n);
    u = {std::fma(-mn_, n.x, m.x), std::fma(-mn_, n.y, m.y), std::fma(-mn_, n.z, m.z)};
    double l = std::sqrt(dot3(u, u));
    if (l < 1e-6) {
        int k = 0;
        for (int i = 1; i < 3; i++)
            if (std::fabs(n[i]) < std::fabs(n[k])) k = i;
        const V3 e = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
        u = cross(n, e);
        l = std::sqrt(dot3(u, u));
    }
    u = scale_inv(u, l);
    v = cross(u, n);
}

double plane_spacing(int lattice, const long h[3], double a0)
{
    const double norm = std::sqrt((double)(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]));
    bool full;
    if (lattice == DISCO_LATTICE_FCC_GEO)
        full = (std::labs(h[0]) % 2) && (std::labs(h[1]) % 2) && (std::labs(h[2]) % 2);
    else
        full = std::labs(h[0] + h[1] + h[2]) % 2 == 0;
    return full ? a0 / norm : a0 / (2.0 * norm);
}

// Is m (in units of a0/2) a translation of the lattice?  Integer components
// within 1e-9; FCC: even sum; BCC: all of one parity (Code/lattice.hpp:159-168).
bool lattice_translation(int lattice, const double m[3])
{
    long h[3];
    for (int i = 0; i < 3; i++) {
        h[i] = std::lround(m[i]);
        if (std::fabs(m[i] - (double)h[i]) > 1e-9) return false;
    }
    if (lattice == DISCO_LATTICE_FCC_GEO) return std::labs(h[0] + h[1] + h[2]) % 2 == 0;
    return (std::labs(h[0]) % 2 == std::labs(h[1]) % 2) &&
           (std::labs(h[1]) % 2 == std::labs(h[2]) % 2);
}

// A loop as the reference's constructors leave it.
struct Loop {
    int shape = 0;
    V3 c{}, mh{}, nh{}, u{}, v{};
    double bmag = 0, radius = 0, half_m = 0, half_p = 0, r_corner = 0;
    std::vector<double> pu, pv, qu, qv, nu, nv;
};

bool has_axis(const double* a) { return a[0] != 0.0 || a[1] != 0.0 || a[2] != 0.0; }

// Polygon part of the constructor (Code/loop.cpp:178-244): vertices on (u, v),
// counter-clockwise, strict convexity, corner clamp, inward normals, inner
// polygon Q.
bool prepare_polygon(const disco_loop_shape* s, Loop& L)
{
    const int N = s->nvert;
    if (N < 3 || !s->vert) return false;
    for (int i = 0; i < N; i++) {
        const V3 x = v3(s->vert + 3 * i);
        L.pu.push_back(dot3(x, L.u));
        L.pv.push_back(dot3(x, L.v));
    }
    double area2 = 0.0;
    for (int i = 0; i < N; i++) {
        const int j = (i + 1) % N;
        area2 += std::fma(L.pu[i], L.pv[j], -(L.pu[j] * L.pv[i]));
    }
    if (std::fabs(area2) < 1e-12) return false;
    if (area2 < 0.0) {
        std::reverse(L.pu.begin(), L.pu.end());
        std::reverse(L.pv.begin(), L.pv.end());
    }
    std::vector<double> du(N), dv(N), len(N), cot_half(N);
    for (int i = 0; i < N; i++) {
        const int j = (i + 1) % N;
        const double eu = L.pu[j] - L.pu[i], ev = L.pv[j] - L.pv[i];
        len[i] = std::sqrt(std::fma(eu, eu, ev * ev));
        if (len[i] < 1e-9) return false;
        du[i] = eu / len[i];
        dv[i] = ev / len[i];
    }
    double turn = 0.0;
    for (int i = 0; i < N; i++) {
        const int k = (i + N - 1) % N;
        const double sn = std::fma(du[k], dv[i], -(dv[k] * du[i]));
        const double cs = std::fma(du[k], du[i], dv[k] * dv[i]);
        if (sn < 1e-9) return false;                       // not strictly convex
        turn += std::atan2(sn, cs);
        cot_half[i] = (1.0 - cs) / sn;
    }
    if (std::fabs(turn - 2.0 * kPiGeo) > 1e-6) return false;   // self-intersecting
    double r_max = 1e300;
    for (int i = 0; i < N; i++) r_max = mn(r_max, len[i] / (cot_half[i] + cot_half[(i + 1) % N]));
    double r = s->corner_r;
    if (r < 0.0) return false;
    if (r > r_max) r = r_max;
    L.r_corner = r;
    L.nu.resize(N);
    L.nv.resize(N);
    for (int i = 0; i < N; i++) {
        L.nu[i] = -dv[i];
        L.nv[i] = du[i];
    }
    L.qu.resize(N);
    L.qv.resize(N);
    for (int i = 0; i < N; i++) {
        const int k = (i + N - 1) % N;
        const double n1u = -dv[k], n1v = du[k], n2u = -dv[i], n2v = du[i];
        const double f = L.r_corner / std::fma(n1v, n2v, std::fma(n1u, n2u, 1.0));
        L.qu[i] = std::fma(f, n1u + n2u, L.pu[i]);
        L.qv[i] = std::fma(f, n1v + n2v, L.pv[i]);
    }
    for (int i = 0; i < N; i++) {
        L.half_m = mx(L.half_m, std::fabs(L.pu[i]));
        L.half_p = mx(L.half_p, std::fabs(L.pv[i]));
    }
    return true;
}

bool prepare(const disco_loop_shape* s, Loop& L)
{
    if (!s || s->shape < DISCO_SHAPE_CIRCLE || s->shape > DISCO_SHAPE_PARALLELOGRAM) return false;
    L.shape = s->shape;
    L.c = v3(s->center);
    // initCommon (Code/loop.cpp:13-35)
    const V3 m = v3(s->m), n = v3(s->n);
    const double norm_m = std::sqrt(dot3(m, m));
    const double norm_n = std::sqrt(dot3(n, n));
    if (!(norm_m > 0.0) || !(norm_n > 0.0)) return false;
    L.mh = scale_inv(m, norm_m);
    L.nh = scale_inv(n, norm_n);
    L.bmag = 1.0 * s->a0 * 0.5 * norm_m;
    switch (s->shape) {
        case DISCO_SHAPE_CIRCLE:
            if (!(s->radius > 0.0)) return false;
            L.radius = s->radius;
            default_axes(L.mh, L.nh, L.u, L.v);
            return true;
        case DISCO_SHAPE_RECT: {
            L.half_m = s->half_m;
            L.half_p = s->half_p;
            if (has_axis(s->axis)) {
                const V3 sd = v3(s->axis);
                L.u = scale_inv(sd, std::sqrt(dot3(sd, sd)));
                L.v = cross(L.u, L.nh);
            } else {
                default_axes(L.mh, L.nh, L.u, L.v);
            }
            const double max_corner = mn(L.half_m, L.half_p);
            L.r_corner = s->corner_r > max_corner ? max_corner : s->corner_r;
            return true;
        }
        case DISCO_SHAPE_PARALLELOGRAM: {
            L.half_m = s->half_m;
            L.half_p = s->half_p;
            const V3 d1 = v3(s->dir1), d2 = v3(s->dir2);
            L.u = scale_inv(d1, std::sqrt(dot3(d1, d1)));
            L.v = scale_inv(d2, std::sqrt(dot3(d2, d2)));
            return true;
        }
        default:
            if (has_axis(s->axis)) {
                const V3 a = v3(s->axis);
                const double an = dot3(a, L.nh);
                const V3 w = {std::fma(-an, L.nh.x, a.x), std::fma(-an, L.nh.y, a.y),
                              std::fma(-an, L.nh.z, a.z)};
                L.u = scale_inv(w, std::sqrt(dot3(w, w)));
                L.v = cross(L.u, L.nh);
            } else {
                default_axes(L.mh, L.nh, L.u, L.v);
            }
            return prepare_polygon(s, L);
    }
}

// chords for an arc of angle th, radius r, sagitta <= eps (Code/loop.cpp:478-484)
int arc_chords(double th, double r, double eps)
{
    if (r <= 0.0 || th <= 0.0) return 0;
    const double c = 1.0 - eps / r;
    const double step = c <= -1.0 ? kPiGeo : 2.0 * std::acos(mx(c, -1.0));
    const int k = (int)std::ceil(th / step - 1e-12);
    return k > 1 ? k : 1;
}

// Code/loop.cpp:489-527: the 2D contour of a (rounded) convex polygon, CCW
void rounded_polygon(const std::vector<double>& pu, const std::vector<double>& pv, double r,
                     double eps, std::vector<double>& ou, std::vector<double>& ov)
{
    const int N = (int)pu.size();
    auto push = [&](double a, double b) {
        if (!ou.empty() && std::fabs(ou.back() - a) < 1e-12 && std::fabs(ov.back() - b) < 1e-12)
            return;
        ou.push_back(a);
        ov.push_back(b);
    };
    if (r <= 0.0) {
        for (int i = 0; i < N; i++) push(pu[i], pv[i]);
    } else {
        std::vector<double> nu(N), nv(N);   // outward unit normals
        for (int i = 0; i < N; i++) {
            const int j = (i + 1) % N;
            const double eu = pu[j] - pu[i], ev = pv[j] - pv[i];
            const double l = std::sqrt(std::fma(eu, eu, ev * ev));
            nu[i] = ev / l;
            nv[i] = -eu / l;
        }
        for (int i = 0; i < N; i++) {
            const int k = (i + N - 1) % N;
            const double f = r / std::fma(nv[k], nv[i], std::fma(nu[k], nu[i], 1.0));
            const double qu = std::fma(-f, nu[k] + nu[i], pu[i]);
            const double qv = std::fma(-f, nv[k] + nv[i], pv[i]);
            const double a0 = std::atan2(nv[k], nu[k]);
            double th = std::atan2(nv[i], nu[i]) - a0;
            while (th < 0.0) th += 2.0 * kPiGeo;
            while (th >= 2.0 * kPiGeo) th -= 2.0 * kPiGeo;
            const int nc = arc_chords(th, r, eps);
            for (int c = 0; c <= nc; c++) {
                const double a = a0 + th * c / (nc > 1 ? nc : 1);
                double sa, ca;
                ::sincos(a, &sa, &ca);
                push(std::fma(r, ca, qu), std::fma(r, sa, qv));
            }
        }
    }
    if (ou.size() > 1 && std::fabs(ou.back() - ou[0]) < 1e-12 &&
        std::fabs(ov.back() - ov[0]) < 1e-12) {
        ou.pop_back();
        ov.pop_back();
    }
}

// contour in plane coordinates (a along u, b along v); nseg > 0 forces the
// chord count of a circle
void plane_coords(const Loop& L, int nseg, double eps, std::vector<double>& a,
                  std::vector<double>& b)
{
    a.clear();
    b.clear();
    if (L.shape == DISCO_SHAPE_CIRCLE) {
        int n = nseg;
        if (n <= 0) {
            const int k = arc_chords(2.0 * kPiGeo, L.radius, eps);
            n = k > 8 ? k : 8;
        }
        const double rp = L.radius * std::sqrt(2.0 * kPiGeo / (n * std::sin(2.0 * kPiGeo / n)));
        for (int k = 0; k < n; k++) {
            double st, ct;
            ::sincos(2.0 * kPiGeo * k / n, &st, &ct);
            a.push_back(rp * ct);
            b.push_back(rp * st);
        }
        return;
    }
    if (L.shape == DISCO_SHAPE_RECT || L.shape == DISCO_SHAPE_PARALLELOGRAM) {
        const std::vector<double> pu = {L.half_m, -L.half_m, -L.half_m, L.half_m};
        const std::vector<double> pv = {L.half_p, L.half_p, -L.half_p, -L.half_p};
        if (L.shape == DISCO_SHAPE_PARALLELOGRAM) {
            a = pu;
            b = pv;
            return;
        }
        rounded_polygon(pu, pv, L.r_corner, eps, a, b);
        return;
    }
    rounded_polygon(L.pu, L.pv, L.r_corner, eps, a, b);
}

// in-plane point to space: u a + v b + c (Code/loop.cpp:564-566, compress = 1)
V3 to_space(const Loop& L, double a, double b)
{
    return {std::fma(L.u.x, a, L.v.x * b) + L.c.x, std::fma(L.u.y, a, L.v.y * b) + L.c.y,
            std::fma(L.u.z, a, L.v.z * b) + L.c.z};
}

std::vector<V3> polyline(const Loop& L, double eps, int nseg)
{
    std::vector<double> a, b;
    plane_coords(L, nseg, eps, a, b);
    const size_t N = a.size();
    std::vector<V3> out(N);
    for (size_t k = 0; k < N; k++) out[k] = to_space(L, a[k], b[k]);
    double s = 0.0;
    for (size_t k = 0; k < N; k++) {
        const V3& p = out[k];
        const V3& q = out[(k + 1) % N];
        const V3 r1 = {p.x - L.c.x, p.y - L.c.y, p.z - L.c.z};
        const V3 r2 = {q.x - L.c.x, q.y - L.c.y, q.z - L.c.z};
        const V3 cr = cross(r1, r2);
        s += std::fma(cr.z, L.nh.z, std::fma(cr.x, L.nh.x, cr.y * L.nh.y));
    }
    if (s < 0.0) std::reverse(out.begin(), out.end());
    return out;
}

// sharp contour with each side cut into collinear pieces, nseg_total in all
// (a workload construct of this producer, not of the reference)
std::vector<V3> subdivided(const Loop& L, int nseg_total)
{
    const std::vector<V3> P = polyline(L, 0.01, 0);
    const int ns = (int)P.size();
    std::vector<V3> out;
    for (int i = 0; i < ns; i++) {
        const V3 A = P[i], B = P[(i + 1) % ns];
        const int per = nseg_total / ns + (i < nseg_total % ns ? 1 : 0);
        for (int j = 0; j < per; j++) {
            const double t = (double)j / per;
            out.push_back({A.x + t * (B.x - A.x), A.y + t * (B.y - A.y), A.z + t * (B.z - A.z)});
        }
    }
    return out;
}

// Code/loop.cpp:281-316
bool inside_uv(const Loop& L, double u, double v)
{
    const double eps = 1e-9;
    if (L.shape == DISCO_SHAPE_CIRCLE) return std::sqrt(std::fma(u, u, v * v)) < L.radius - eps;
    if (L.shape == DISCO_SHAPE_POLYGON) {
        const size_t N = L.qu.size();
        bool in_q = true;
        for (size_t i = 0; i < N && in_q; i++)
            if (std::fma(L.nu[i], u - L.pu[i], L.nv[i] * (v - L.pv[i])) < L.r_corner - eps)
                in_q = false;
        if (in_q) return true;
        if (L.r_corner <= 0.0) return false;
        double d2 = 1e300;
        for (size_t i = 0; i < N; i++) {
            const size_t j = (i + 1) % N;
            const double eu = L.qu[j] - L.qu[i], ev = L.qv[j] - L.qv[i];
            const double l2 = std::fma(eu, eu, ev * ev);
            double t = l2 > 0.0 ? std::fma(u - L.qu[i], eu, (v - L.qv[i]) * ev) / l2 : 0.0;
            t = mn(mx(t, 0.0), 1.0);
            const double wu = std::fma(-t, eu, u - L.qu[i]);
            const double wv = std::fma(-t, ev, v - L.qv[i]);
            d2 = mn(d2, std::fma(wu, wu, wv * wv));
        }
        return d2 <= std::fma(L.r_corner, L.r_corner, eps);
    }
    const double du = mx(std::fabs(u) - mx(L.half_m - L.r_corner, 0.0), 0.0);
    const double dv = mx(std::fabs(v) - mx(L.half_p - L.r_corner, 0.0), 0.0);
    return std::fma(du, du, dv * dv) <= std::fma(L.r_corner, L.r_corner, eps);
}

double surface(const Loop& L, double da)
{
    double ds = da * da;
    if (L.shape == DISCO_SHAPE_PARALLELOGRAM) {
        const V3 cr = cross(L.u, L.v);
        ds *= std::sqrt(dot3(cr, cr));
    }
    return ds;
}

template <typename P>
int emit(const std::vector<V3>& pts, P* x, P* y, P* z, int cap)
{
    const int n = (int)pts.size();
    if (n <= cap && x && y && z)
        for (int i = 0; i < n; i++) { x[i] = pts[i].x; y[i] = pts[i].y; z[i] = pts[i].z; }
    return n;
}

void put3(double* o, const V3& a)
{
    o[0] = a.x;
    o[1] = a.y;
    o[2] = a.z;
}

}  // namespace

extern "C" {

int disco_geo_regular_polygon(int sides, double R, const double m[3], const double n[3],
                              double* out)
{
    if (sides < 3 || !m || !n || !out) return DISCO_GEO_ERR_ARG;
    const V3 mm = v3(m), nn = v3(n);
    V3 u, v;
    default_axes(scale_inv(mm, std::sqrt(dot3(mm, mm))), scale_inv(nn, std::sqrt(dot3(nn, nn))),
                 u, v);
    for (int k = 0; k < sides; k++) {
        const double th = 2.0 * kPiGeo * k / sides;
        double st, ct;
        ::sincos(th, &st, &ct);
        out[3 * k + 0] = R * std::fma(ct, u.x, st * v.x);
        out[3 * k + 1] = R * std::fma(ct, u.y, st * v.y);
        out[3 * k + 2] = R * std::fma(ct, u.z, st * v.z);
    }
    return sides;
}

int disco_geo_snap_mid_plane(const double c[3], const double n[3], double a0, int nx, int ny,
                             int nz, int lattice, double out[3], double* spacing)
{
    if (!c || !n || !out) return DISCO_GEO_ERR_ARG;
    long h[3];
    for (int i = 0; i < 3; i++) {
        h[i] = std::lround(n[i]);
        if (std::fabs(n[i] - (double)h[i]) > 1e-9) return DISCO_GEO_ERR_ARG;   // Miller indices
    }
    long g = 0;
    for (int i = 0; i < 3; i++) {
        long a = std::labs(h[i]);
        while (a) { const long t = g % a; g = a; a = t; }
    }
    if (g == 0) return DISCO_GEO_ERR_ARG;
    for (int i = 0; i < 3; i++) h[i] /= g;
    const double norm = std::sqrt((double)(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]));
    const V3 nh = {h[0] / norm, h[1] / norm, h[2] / norm};
    const double sp = plane_spacing(lattice, h, a0);
    const V3 o = {-0.5 * a0 * nx, -0.5 * a0 * ny, -0.5 * a0 * nz};
    const double s0 = dot3(o, nh);
    const double s = dot3(v3(c), nh);
    const double s_mid = std::fma(std::floor((s - s0) / sp + 1e-9) + 0.5, sp, s0);
    out[0] = std::fma(s_mid - s, nh.x, c[0]);
    out[1] = std::fma(s_mid - s, nh.y, c[1]);
    out[2] = std::fma(s_mid - s, nh.z, c[2]);
    if (spacing) *spacing = sp;
    return 1;
}

int disco_geo_polyline(const disco_loop_shape* s, double eps, int nseg, int subdivide,
                       double* vx, double* vy, double* vz, int cap)
{
    Loop L;
    if (!prepare(s, L)) return DISCO_GEO_ERR_ARG;
    if (subdivide > 0) return emit(subdivided(L, subdivide), vx, vy, vz, cap);
    return emit(polyline(L, eps, nseg), vx, vy, vz, cap);
}

int disco_geo_loopons(const disco_loop_shape* s, double da, double* x, double* y, double* z,
                      int cap)
{
    Loop L;
    if (!prepare(s, L) || !(da > 0.0)) return DISCO_GEO_ERR_ARG;
    const double ext_m = L.shape == DISCO_SHAPE_CIRCLE ? L.radius : L.half_m;
    const double ext_p = L.shape == DISCO_SHAPE_CIRCLE ? L.radius : L.half_p;
    const int sm = (int)std::ceil(ext_m / da) + 1;
    const int sp = (int)std::ceil(ext_p / da) + 1;
    std::vector<V3> pts;
    for (int i = -sm; i <= sm; i++)
        for (int j = -sp; j <= sp; j++) {
            const double U = i * da, V = j * da;
            if (inside_uv(L, U, V)) pts.push_back(to_space(L, U, V));
        }
    return emit(pts, x, y, z, cap);
}

int disco_geo_phys(const disco_loop_shape* s, double da, double phys[10])
{
    Loop L;
    if (!prepare(s, L) || !phys) return DISCO_GEO_ERR_ARG;
    const double p[10] = {L.bmag, L.bmag, L.bmag, L.mh.x, L.mh.y, L.mh.z,
                          L.nh.x, L.nh.y, L.nh.z, da > 0 ? surface(L, da) : 0.0};
    std::memcpy(phys, p, sizeof(p));
    return 10;
}

int disco_geo_loop_state(const disco_loop_shape* s, disco_loop_state* out)
{
    Loop L;
    if (!out || !prepare(s, L)) return DISCO_GEO_ERR_ARG;
    if (L.pu.size() > DISCO_MAX_STATE_VERT) return DISCO_GEO_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    out->shape = L.shape;
    out->nvert = (int)L.pu.size();
    put3(out->center, L.c);
    put3(out->m, L.mh);
    put3(out->n, L.nh);
    put3(out->u, L.u);
    put3(out->v, L.v);
    out->bmag = L.bmag;
    out->radius = L.radius;
    out->half_m = L.half_m;
    out->half_p = L.half_p;
    out->r_corner = L.r_corner;
    for (size_t i = 0; i < L.pu.size(); i++) {
        out->pu[i] = L.pu[i];
        out->pv[i] = L.pv[i];
        out->qu[i] = L.qu[i];
        out->qv[i] = L.qv[i];
        out->nu[i] = L.nu[i];
        out->nv[i] = L.nv[i];
    }
    return 1;
}

long long disco_geo_polylines(const disco_loop_shape* shapes, int nloops, double eps,
                              const int* nseg, const int* subdivide, long long* offsets,
                              double* vx, double* vy, double* vz, long long cap)
{
    if (nloops < 0 || (nloops > 0 && (!shapes || !offsets))) return DISCO_GEO_ERR_ARG;
    std::vector<std::vector<V3>> all(nloops);
    long long total = 0;
    offsets[0] = 0;
    for (int l = 0; l < nloops; l++) {
        Loop L;
        if (!prepare(&shapes[l], L)) return DISCO_GEO_ERR_ARG;
        const int sub = subdivide ? subdivide[l] : 0;
        all[l] = sub > 0 ? subdivided(L, sub) : polyline(L, eps, nseg ? nseg[l] : 0);
        total += (long long)all[l].size();
        offsets[l + 1] = total;
    }
    if (total <= cap && vx && vy && vz)
        for (int l = 0; l < nloops; l++) {
            long long o = offsets[l];
            for (const V3& p : all[l]) { vx[o] = p.x; vy[o] = p.y; vz[o] = p.z; o++; }
        }
    return total;
}

int disco_geo_classify_burgers(int lattice, const double m[3], const double n[3])
{
    if (!m || !n) return DISCO_GEO_ERR_ARG;
    if (lattice_translation(lattice, m)) return DISCO_BURGERS_PERFECT;
    if (lattice != DISCO_LATTICE_FCC_GEO) return DISCO_BURGERS_OTHER;
    long k[3];   // 3 m: integers for the 1/3 fractions of the partials
    for (int i = 0; i < 3; i++) {
        k[i] = std::lround(3.0 * m[i]);
        if (std::fabs(std::fma(3.0, m[i], -(double)k[i])) > 1e-6) return DISCO_BURGERS_OTHER;
    }
    const long a = std::labs(k[0]), b = std::labs(k[1]), c = std::labs(k[2]);
    if (a == 2 && b == 2 && c == 2) {   // 2/3<111>: a Frank partial when n is along m
        const double x[3] = {std::fma(m[1], n[2], -(m[2] * n[1])),
                             std::fma(m[2], n[0], -(m[0] * n[2])),
                             std::fma(m[0], n[1], -(m[1] * n[0]))};
        const double cn = std::sqrt(std::fma(x[2], x[2], std::fma(x[0], x[0], x[1] * x[1])));
        const double nn = std::sqrt(std::fma(n[2], n[2], std::fma(n[0], n[0], n[1] * n[1])));
        return cn < 1e-6 * nn ? DISCO_BURGERS_FRANK : DISCO_BURGERS_OTHER;
    }
    const long lo = std::min(a, std::min(b, c)), hi = std::max(a, std::max(b, c));
    if (lo == 1 && hi == 2 && a + b + c == 4) {   // 1/3<112> in a {111} plane, m.n = 0
        const bool n111 = std::fabs(n[0]) > 0.0 &&
                          std::fabs(std::fabs(n[0]) - std::fabs(n[1])) < 1e-9 &&
                          std::fabs(std::fabs(n[0]) - std::fabs(n[2])) < 1e-9;
        const double mn = std::fma(m[2], n[2], std::fma(m[0], n[0], m[1] * n[1]));
        return n111 && std::fabs(mn) < 1e-6 ? DISCO_BURGERS_SHOCKLEY : DISCO_BURGERS_OTHER;
    }
    return DISCO_BURGERS_OTHER;
}

void disco_geo_frank_copy_shift(const double v[3], const double nhat[3], double a0,
                                double out[3])
{
    // the 24 vectors a0/6 <112>: the 2 sits on each axis in turn (cyclic),
    // each of the 8 sign patterns; the first one in the plane whose sum with v
    // is not a translation wins
    for (int p = 0; p < 3; p++)
        for (int sg = 0; sg < 8; sg++) {
            double s[3];
            for (int i = 0; i < 3; i++) {
                const double sign = (sg >> i) & 1 ? -1.0 : 1.0;
                const double base = i == 2 ? 2.0 : 1.0;
                s[(p + i) % 3] = sign * base * a0 / 6.0;
            }
            const double sn = std::fma(s[2], nhat[2], std::fma(s[0], nhat[0], s[1] * nhat[1]));
            if (std::fabs(sn) > 1e-9 * a0) continue;
            const double w[3] = {2.0 * (v[0] + s[0]) / a0, 2.0 * (v[1] + s[1]) / a0,
                                 2.0 * (v[2] + s[2]) / a0};
            if (lattice_translation(DISCO_LATTICE_FCC_GEO, w)) continue;
            for (int i = 0; i < 3; i++) out[i] = v[i] + s[i];
            return;
        }
    for (int i = 0; i < 3; i++) out[i] = v[i];
}

void disco_geo_copy_shifts(const double b[3], int frank, const double n[3], const double* R,
                           double a0, double shift[6])
{
    // rotateToCrystal / rotateToLab (Code/lattice.hpp:64-78): R^T v, R c
    auto to_crystal = [R](const double* v, double* c) {
        for (int i = 0; i < 3; i++)
            c[i] = std::fma(R[6 + i], v[2], std::fma(R[i], v[0], R[3 + i] * v[1]));
    };
    auto to_lab = [R](const double* c, double* r) {
        for (int i = 0; i < 3; i++)
            r[i] = std::fma(R[3 * i + 2], c[2], std::fma(R[3 * i], c[0], R[3 * i + 1] * c[1]));
    };
    for (int sd = 0; sd < 2; sd++) {
        const double sign = sd == 0 ? 1.0 : -1.0;
        const double v[3] = {sign * b[0], sign * b[1], sign * b[2]};
        double* out = shift + 3 * sd;
        if (!frank) {
            for (int i = 0; i < 3; i++) out[i] = v[i];
            continue;
        }
        double nh[3], vc[3], w[3];
        if (R) {
            to_crystal(n, nh);
            to_crystal(v, vc);
        } else {
            for (int i = 0; i < 3; i++) { nh[i] = n[i]; vc[i] = v[i]; }
        }
        disco_geo_frank_copy_shift(vc, nh, a0, w);
        if (R)
            to_lab(w, out);
        else
            for (int i = 0; i < 3; i++) out[i] = w[i];
    }
}

}  // extern "C"
