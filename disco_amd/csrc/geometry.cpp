// geometry.cpp -- host producer of loop inputs (include/disco_geometry.h).
//
// Follows the reference's loop code formula by formula and in its operation
// order (every product and sum rounded on its own: contraction is off in this
// file), so the polylines, loopon grids and snapped centres equal the
// reference's: regularPolygon (Code/loop.cpp:585-608), defaultAxes (:40-57),
// snapToMidPlane (:342-374) with latticePlaneSpacing (Code/lattice.hpp:146-153),
// loop::polyline (:536-580) and generateLoopons / insideUV (:253-316).
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

double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(const V3& a, const V3& b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
V3 unit(const V3& v)
{
    const double l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return {v.x / l, v.y / l, v.z / l};
}
V3 v3(const double* p) { return {p[0], p[1], p[2]}; }

// in-plane axes (u, v) for unit m, n (Code/loop.cpp:40-57)
void default_axes(const V3& m, const V3& n, V3& u, V3& v)
{
    const double mn = m.x * n.x + m.y * n.y + m.z * n.z;
    u = {m.x - mn * n.x, m.y - mn * n.y, m.z - mn * n.z};
    double l = std::sqrt(u.x * u.x + u.y * u.y + u.z * u.z);
    if (l < 1e-6) {
        int k = 0;
        for (int i = 1; i < 3; i++)
            if (std::fabs(n[i]) < std::fabs(n[k])) k = i;
        const V3 e = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
        u = cross(n, e);
        l = std::sqrt(u.x * u.x + u.y * u.y + u.z * u.z);
    }
    u = {u.x / l, u.y / l, u.z / l};
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

// A loop prepared as the reference prepares it: unit m and n, |b|, the axes,
// and for a polygon its in-plane coordinates, counter-clockwise.
struct Loop {
    int shape = 0;
    V3 c{}, mh{}, nh{}, u{}, v{};
    double bmag = 0, radius = 0, half_m = 0, half_p = 0;
    std::vector<double> pu, pv;
};

bool prepare(const disco_loop_shape* s, Loop& L)
{
    if (!s || s->shape < DISCO_SHAPE_CIRCLE || s->shape > DISCO_SHAPE_PARALLELOGRAM) return false;
    L.shape = s->shape;
    L.c = v3(s->center);
    const double normm = std::sqrt(s->m[0] * s->m[0] + s->m[1] * s->m[1] + s->m[2] * s->m[2]);
    if (!(normm > 0.0)) return false;
    L.mh = {s->m[0] / normm, s->m[1] / normm, s->m[2] / normm};
    const V3 n = v3(s->n);
    if (!(dot(n, n) > 0.0)) return false;
    L.nh = unit(n);
    L.bmag = s->a0 * 0.5 * normm;
    L.radius = s->radius;
    L.half_m = s->half_m;
    L.half_p = s->half_p;
    if (s->shape == DISCO_SHAPE_PARALLELOGRAM) {
        L.u = unit(v3(s->dir1));
        L.v = unit(v3(s->dir2));
    } else {
        default_axes(L.mh, L.nh, L.u, L.v);
    }
    if (s->shape == DISCO_SHAPE_CIRCLE && !(s->radius > 0.0)) return false;
    if (s->shape == DISCO_SHAPE_POLYGON) {
        if (s->nvert < 3 || !s->vert) return false;
        const int N = s->nvert;
        L.pu.resize(N);
        L.pv.resize(N);
        for (int i = 0; i < N; i++) {
            const V3 x = v3(s->vert + 3 * i);
            L.pu[i] = x.x * L.u.x + x.y * L.u.y + x.z * L.u.z;
            L.pv[i] = x.x * L.v.x + x.y * L.v.y + x.z * L.v.z;
        }
        double a2 = 0.0;
        for (int i = 0; i < N; i++)
            a2 += L.pu[i] * L.pv[(i + 1) % N] - L.pu[(i + 1) % N] * L.pv[i];
        if (a2 < 0.0) {
            std::vector<double> ru(L.pu.rbegin(), L.pu.rend()), rv(L.pv.rbegin(), L.pv.rend());
            L.pu.swap(ru);
            L.pv.swap(rv);
        }
        double hm = 0.0, hp = 0.0;
        for (int i = 0; i < N; i++) {
            hm = std::fmax(hm, std::fabs(L.pu[i]));
            hp = std::fmax(hp, std::fabs(L.pv[i]));
        }
        L.half_m = hm;
        L.half_p = hp;
    }
    return true;
}

// contour in the plane coordinates (a along u, b along v)
void plane_coords(const Loop& L, int nseg, double eps, std::vector<double>& a,
                  std::vector<double>& b)
{
    a.clear();
    b.clear();
    if (L.shape == DISCO_SHAPE_CIRCLE) {
        if (nseg <= 0) {
            const double c = 1.0 - eps / L.radius;
            const double step = c <= -1.0 ? kPiGeo : 2.0 * std::acos(std::fmax(c, -1.0));
            const int n = (int)std::ceil(2.0 * kPiGeo / step - 1e-12);
            nseg = n > 8 ? n : 8;
        }
        const double rp = L.radius * std::sqrt(2.0 * kPiGeo / (nseg * std::sin(2.0 * kPiGeo / nseg)));
        for (int k = 0; k < nseg; k++) {
            a.push_back(rp * std::cos(2.0 * kPiGeo * k / nseg));
            b.push_back(rp * std::sin(2.0 * kPiGeo * k / nseg));
        }
        return;
    }
    if (L.shape == DISCO_SHAPE_RECT || L.shape == DISCO_SHAPE_PARALLELOGRAM) {
        a = {L.half_m, -L.half_m, -L.half_m, L.half_m};
        b = {L.half_p, L.half_p, -L.half_p, -L.half_p};
        return;
    }
    for (size_t i = 0; i < L.pu.size(); i++) {
        if (!a.empty() && std::fabs(a.back() - L.pu[i]) < 1e-12 &&
            std::fabs(b.back() - L.pv[i]) < 1e-12)
            continue;
        a.push_back(L.pu[i]);
        b.push_back(L.pv[i]);
    }
    if (a.size() > 1 && std::fabs(a.back() - a[0]) < 1e-12 && std::fabs(b.back() - b[0]) < 1e-12) {
        a.pop_back();
        b.pop_back();
    }
}

std::vector<V3> polyline(const Loop& L, double eps, int nseg)
{
    std::vector<double> a, b;
    plane_coords(L, nseg, eps, a, b);
    const size_t N = a.size();
    std::vector<V3> out(N);
    for (size_t k = 0; k < N; k++)
        out[k] = {(L.u.x * a[k] + L.v.x * b[k]) + L.c.x, (L.u.y * a[k] + L.v.y * b[k]) + L.c.y,
                  (L.u.z * a[k] + L.v.z * b[k]) + L.c.z};
    double s = 0.0;
    for (size_t k = 0; k < N; k++) {
        const V3& p = out[k];
        const V3& q = out[(k + 1) % N];
        const V3 cr = cross({p.x - L.c.x, p.y - L.c.y, p.z - L.c.z},
                            {q.x - L.c.x, q.y - L.c.y, q.z - L.c.z});
        s += cr.x * L.nh.x + cr.y * L.nh.y + cr.z * L.nh.z;
    }
    if (s < 0.0) std::vector<V3>(out.rbegin(), out.rend()).swap(out);
    return out;
}

// sharp contour with each side cut into collinear pieces, nseg_total in all
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

bool inside_uv(const Loop& L, double u, double v)
{
    const double eps = 1e-9;
    if (L.shape == DISCO_SHAPE_CIRCLE) return std::sqrt(u * u + v * v) < L.radius - eps;
    if (L.shape == DISCO_SHAPE_RECT || L.shape == DISCO_SHAPE_PARALLELOGRAM) {
        const double du = std::fmax(std::fabs(u) - std::fmax(L.half_m, 0.0), 0.0);
        const double dv = std::fmax(std::fabs(v) - std::fmax(L.half_p, 0.0), 0.0);
        return du * du + dv * dv <= eps;
    }
    const size_t N = L.pu.size();
    for (size_t i = 0; i < N; i++) {
        const size_t j = (i + 1) % N;
        const double eu = L.pu[j] - L.pu[i], ev = L.pv[j] - L.pv[i];
        const double len = std::sqrt(eu * eu + ev * ev);
        const double nu = -ev / len, nv = eu / len;
        if (!(nu * (u - L.pu[i]) + nv * (v - L.pv[i]) >= -eps)) return false;
    }
    return true;
}

double surface(const Loop& L, double da)
{
    double ds = da * da;
    if (L.shape == DISCO_SHAPE_PARALLELOGRAM) {
        const V3 cr = cross(L.u, L.v);
        ds *= std::sqrt(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
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

}  // namespace

extern "C" {

int disco_geo_regular_polygon(int sides, double R, const double m[3], const double n[3],
                              double* out)
{
    if (sides < 3 || !m || !n || !out) return DISCO_GEO_ERR_ARG;
    V3 u, v;
    default_axes(unit(v3(m)), unit(v3(n)), u, v);
    for (int k = 0; k < sides; k++) {
        const double th = 2.0 * kPiGeo * k / sides;
        const double ct = std::cos(th), st = std::sin(th);
        out[3 * k + 0] = R * (ct * u.x + st * v.x);
        out[3 * k + 1] = R * (ct * u.y + st * v.y);
        out[3 * k + 2] = R * (ct * u.z + st * v.z);
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
    const double s0 = o.x * nh.x + o.y * nh.y + o.z * nh.z;
    const double s = c[0] * nh.x + c[1] * nh.y + c[2] * nh.z;
    const double s_mid = s0 + (std::floor((s - s0) / sp + 1e-9) + 0.5) * sp;
    out[0] = c[0] + (s_mid - s) * nh.x;
    out[1] = c[1] + (s_mid - s) * nh.y;
    out[2] = c[2] + (s_mid - s) * nh.z;
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
            if (!inside_uv(L, U, V)) continue;
            pts.push_back({(L.u.x * U + L.v.x * V) + L.c.x, (L.u.y * U + L.v.y * V) + L.c.y,
                           (L.u.z * U + L.v.z * V) + L.c.z});
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

}  // extern "C"
