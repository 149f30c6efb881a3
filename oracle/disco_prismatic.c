/*
 * disco_prismatic.c -- CPU restatement of DISCO's prismatic slab selection.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker of the device slab
 * selection: only tests/ and bench.py's cpu_baseline leg may load it.  The
 * product path never calls it and has no CPU fallback.
 *
 *   oracle_matter_type     matterType     (Code/prismatic.cpp:12-18)
 *   oracle_prismatic_slab  prismaticSlab  (Code/prismatic.cpp:30-74) with
 *                          minImage (:28-31), loop::contains (Code/loop.cpp:323-330)
 *                          and loop::insideUV (Code/loop.cpp:281-316)
 *
 * The loop is given by its constructed state (disco_loop_state).  Built like
 * the rest of the oracle, with the reference's flags and FP contraction, and
 * written in the reference's expression shapes, so the compiler fuses the same
 * products; the tests pin the selection against the reference's own
 * prismaticSlab compiled from its sources, and against brute force in exact
 * arithmetic on small cases.
 */
#include <math.h>
#include <stddef.h>

#include "disco_geometry.h"

static double max_d(double a, double b) { return a < b ? b : a; }
static double min_d(double a, double b) { return b < a ? b : a; }

/* Code/prismatic.cpp:28-31 */
static double min_image(double r, double L) { return r - L * floor(r / L + 0.5); }

/* Code/loop.cpp:281-316 */
static int inside_uv(const disco_loop_state* L, double u, double v)
{
    const double eps = 1e-9;
    if (L->shape == DISCO_SHAPE_CIRCLE) return sqrt(u * u + v * v) < L->radius - eps;
    if (L->shape == DISCO_SHAPE_POLYGON) {
        const int N = L->nvert;
        int in_q = 1;
        for (int i = 0; i < N && in_q; i++)
            if (L->nu[i] * (u - L->pu[i]) + L->nv[i] * (v - L->pv[i]) < L->r_corner - eps) in_q = 0;
        if (in_q) return 1;
        if (L->r_corner <= 0.) return 0;
        double d2 = 1e300;
        for (int i = 0; i < N; i++) {
            const int j = (i + 1) % N;
            const double eu = L->qu[j] - L->qu[i], ev = L->qv[j] - L->qv[i];
            const double l2 = eu * eu + ev * ev;
            double t = l2 > 0. ? ((u - L->qu[i]) * eu + (v - L->qv[i]) * ev) / l2 : 0.;
            t = min_d(max_d(t, 0.), 1.);
            const double wu = u - L->qu[i] - t * eu, wv = v - L->qv[i] - t * ev;
            d2 = min_d(d2, wu * wu + wv * wv);
        }
        return d2 <= L->r_corner * L->r_corner + eps;
    }
    const double du = max_d(fabs(u) - max_d(L->half_m - L->r_corner, 0.0), 0.0);
    const double dv = max_d(fabs(v) - max_d(L->half_p - L->r_corner, 0.0), 0.0);
    return du * du + dv * dv <= L->r_corner * L->r_corner + eps;
}

/* Code/loop.cpp:323-330 */
static int contains(const disco_loop_state* L, const double rel[3])
{
    const double ru = rel[0] * L->u[0] + rel[1] * L->u[1] + rel[2] * L->u[2];
    const double rv = rel[0] * L->v[0] + rel[1] * L->v[1] + rel[2] * L->v[2];
    const double c = L->u[0] * L->v[0] + L->u[1] * L->v[1] + L->u[2] * L->v[2];
    const double det = 1.0 - c * c;
    return inside_uv(L, (ru - c * rv) / det, (rv - c * ru) / det);
}

/* 0 glide, 1 vacancy, 2 interstitial */
int oracle_matter_type(const disco_loop_state* L)
{
    const double mn = L->m[0] * L->n[0] + L->m[1] * L->n[1] + L->m[2] * L->n[2];
    if (fabs(mn) < 1e-6) return 0;
    return mn > 0. ? 1 : 2;
}

/* info = (type, planes, height, spacing, b0, b1, b2).  Returns the slab size
 * (atoms / side written up to cap), or -1 when |b.n| is not a whole number of
 * plane spacings (where the reference aborts). */
long long oracle_prismatic_slab(const disco_loop_state* L, double spacing, long long n,
                                const double* x, const double* y, const double* z, double lx,
                                double ly, double lz, int* atoms, int* side, long long cap,
                                double* info)
{
    const int type = oracle_matter_type(L);
    for (int i = 0; i < 7; i++) info[i] = 0.;
    info[0] = type;
    if (type == 0) return 0;
    double b[3];
    for (int i = 0; i < 3; i++) b[i] = L->bmag * L->m[i];
    const double height = fabs(b[0] * L->n[0] + b[1] * L->n[1] + b[2] * L->n[2]);
    const double ratio = height / spacing;
    const int planes = (int)lround(ratio);
    info[1] = planes;
    info[2] = height;
    info[3] = spacing;
    for (int i = 0; i < 3; i++) info[4 + i] = b[i];
    if (planes < 1 || fabs(ratio - planes) > 1e-6) return -1;
    const double t_lo = -floor(planes / 2.);
    const double t_hi = ceil(planes / 2.);
    long long count = 0;
    for (long long a = 0; a < n; a++) {
        const double rel[3] = {min_image(x[a] - L->center[0], lx),
                               min_image(y[a] - L->center[1], ly),
                               min_image(z[a] - L->center[2], lz)};
        const double s = rel[0] * L->n[0] + rel[1] * L->n[1] + rel[2] * L->n[2];
        const double t = s / spacing;
        if (t < t_lo || t >= t_hi) continue;
        if (!contains(L, rel)) continue;
        if (count < cap) {
            atoms[count] = (int)a;
            side[count] = s > 0. ? 1 : -1;
        }
        count++;
    }
    return count;
}
