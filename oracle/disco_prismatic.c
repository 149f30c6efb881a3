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
 *   oracle_classify_burgers   classifyBurgers (Code/lattice.hpp:189-212) with
 *                             isLatticeTranslation (:159-168)
 *   oracle_frank_copy_shift   frankCopyShift  (Code/lattice.hpp:239-256)
 *   oracle_copy_shifts        the two copy shifts of an interstitial loop
 *                             (Code/main.cpp:633-650), rotateToCrystal /
 *                             rotateToLab (Code/lattice.hpp:64-78)
 *   oracle_prismatic_apply    the driver's use of the slabs: grain / twin
 *                             filter, vacancy deletion and renumbering, the
 *                             copy list (Code/main.cpp:593-687)
 *   oracle_append_copies      the copies after the sum (Code/main.cpp:733-742)
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
#include <stdlib.h>

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

/* ---- applying the slabs ------------------------------------------------- */

/* Code/lattice.hpp:159-168 */
static int is_translation(int fcc, const double m[3])
{
    long h[3];
    for (int i = 0; i < 3; i++) {
        h[i] = lround(m[i]);
        if (fabs(m[i] - h[i]) > 1e-9) return 0;
    }
    if (fcc) return labs(h[0] + h[1] + h[2]) % 2 == 0;
    return (labs(h[0]) % 2 == labs(h[1]) % 2) && (labs(h[1]) % 2 == labs(h[2]) % 2);
}

/* Code/lattice.hpp:189-212; returns DISCO_BURGERS_* */
int oracle_classify_burgers(int lattice, const double* m, const double* n)
{
    const int fcc = lattice == DISCO_LATTICE_FCC_GEO;
    if (is_translation(fcc, m)) return DISCO_BURGERS_PERFECT;
    if (!fcc) return DISCO_BURGERS_OTHER;
    long k[3];
    for (int i = 0; i < 3; i++) {
        k[i] = lround(3. * m[i]);
        if (fabs(3. * m[i] - k[i]) > 1e-6) return DISCO_BURGERS_OTHER;
    }
    const long a = labs(k[0]), b = labs(k[1]), c = labs(k[2]);
    if (a == 2 && b == 2 && c == 2) {
        const double x[3] = { m[1] * n[2] - m[2] * n[1], m[2] * n[0] - m[0] * n[2],
                              m[0] * n[1] - m[1] * n[0] };
        const double cn = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        return cn < 1e-6 * nn ? DISCO_BURGERS_FRANK : DISCO_BURGERS_OTHER;
    }
    long lo = a < b ? a : b, hi = a < b ? b : a;
    if (c < lo) lo = c;
    if (c > hi) hi = c;
    if (lo == 1 && hi == 2 && a + b + c == 4) {
        const int n111 = fabs(n[0]) > 0. && fabs(fabs(n[0]) - fabs(n[1])) < 1e-9 &&
                         fabs(fabs(n[0]) - fabs(n[2])) < 1e-9;
        const double mn = m[0] * n[0] + m[1] * n[1] + m[2] * n[2];
        return n111 && fabs(mn) < 1e-6 ? DISCO_BURGERS_SHOCKLEY : DISCO_BURGERS_OTHER;
    }
    return DISCO_BURGERS_OTHER;
}

/* Code/lattice.hpp:239-256: the first a0/6<112> vector (axis of the 2 cycled,
 * then the 8 sign patterns) normal to nhat whose sum with v is not an FCC
 * translation */
void oracle_frank_copy_shift(const double* v, const double* nhat, double a0, double* out)
{
    for (int p = 0; p < 3; p++)
        for (int sg = 0; sg < 8; sg++) {
            const double mag[3] = { 1., 1., 2. };
            double s[3], w[3];
            for (int i = 0; i < 3; i++)
                s[(p + i) % 3] = (((sg >> i) & 1) ? -1. : 1.) * mag[i] * a0 / 6.;
            if (fabs(s[0] * nhat[0] + s[1] * nhat[1] + s[2] * nhat[2]) > 1e-9 * a0) continue;
            for (int i = 0; i < 3; i++) w[i] = 2. * (v[i] + s[i]) / a0;
            if (is_translation(1, w)) continue;
            for (int i = 0; i < 3; i++) out[i] = v[i] + s[i];
            return;
        }
    for (int i = 0; i < 3; i++) out[i] = v[i];
}

/* Code/main.cpp:633-650; R row-major (rows: crystal directions of X, Y, Z) or
 * NULL (no rotation) */
void oracle_copy_shifts(const double* b, int frank, const double* n, const double* R, double a0,
                        double* shift)
{
    for (int sd = 0; sd < 2; sd++) {
        const double sign = sd == 0 ? 1. : -1.;
        const double v[3] = { sign * b[0], sign * b[1], sign * b[2] };
        double* out = shift + 3 * sd;
        if (!frank) {
            for (int i = 0; i < 3; i++) out[i] = v[i];
            continue;
        }
        double nh[3], vc[3], w[3];
        for (int i = 0; i < 3; i++) {
            nh[i] = R ? R[i] * n[0] + R[3 + i] * n[1] + R[6 + i] * n[2] : n[i];
            vc[i] = R ? R[i] * v[0] + R[3 + i] * v[1] + R[6 + i] * v[2] : v[i];
        }
        oracle_frank_copy_shift(vc, nh, a0, w);
        for (int i = 0; i < 3; i++)
            out[i] = R ? R[3 * i] * w[0] + R[3 * i + 1] * w[1] + R[3 * i + 2] * w[2] : w[i];
    }
}

/* Code/main.cpp:593-687.  Loop l (type[l]: 0 glide, skipped; 1 vacancy;
 * 2 interstitial) has slab entries off[l] .. off[l+1] of atoms / side, its
 * grain loop_grain[l]
 * (< 0: none) and twin region loop_region[l], and its copy shifts
 * shifts[6l .. 6l+5].  x, y, z (and grain, when not NULL) are compacted in
 * place; the copies (source, shift) go to copy_atom / copy_shift (room for all
 * interstitial slab entries), *ncopies their count.  Returns the new atom
 * count, -1 when out of memory. */
long long oracle_prismatic_apply(long long n, double* x, double* y, double* z, int* grain,
                                 const int* twin, int nloops, const int* type,
                                 const long long* off, const int* atoms, const int* side,
                                 const int* loop_grain, const int* loop_region,
                                 const double* shifts, int* copy_atom, double* copy_shift,
                                 long long* ncopies)
{
    unsigned char* removed = (unsigned char*)calloc(n > 0 ? (size_t)n : 1, 1);
    int* new_index = (int*)malloc((n > 0 ? (size_t)n : 1) * sizeof(int));
    if (!removed || !new_index) { free(removed); free(new_index); return -1; }
    long long nc = 0;
    int any = 0;
    for (int l = 0; l < nloops; l++) {
        if (type[l] == 0) continue;
        any = 1;
        const int gl = loop_grain[l];
        for (long long q = off[l]; q < off[l + 1]; q++) {
            const int a = atoms[q];
            if (gl >= 0 && grain[a] != gl) continue;
            if (twin && twin[a] != loop_region[l]) continue;
            if (type[l] == 1) {
                removed[a] = 1;
            } else {
                const double* sh = shifts + 6 * l + (side[q] > 0 ? 0 : 3);
                copy_atom[nc] = a;
                for (int c = 0; c < 3; c++) copy_shift[3 * nc + c] = sh[c];
                nc++;
            }
        }
    }
    long long k = n;
    if (any) {
        k = 0;
        for (long long a = 0; a < n; a++) {
            if (removed[a]) { new_index[a] = -1; continue; }
            new_index[a] = (int)k;
            x[k] = x[a]; y[k] = y[a]; z[k] = z[a];
            if (grain) grain[k] = grain[a];
            k++;
        }
        long long kept = 0;
        for (long long c = 0; c < nc; c++) {
            if (new_index[copy_atom[c]] < 0) continue;
            copy_atom[kept] = new_index[copy_atom[c]];
            for (int i = 0; i < 3; i++) copy_shift[3 * kept + i] = copy_shift[3 * c + i];
            kept++;
        }
        nc = kept;
    }
    free(removed);
    free(new_index);
    *ncopies = nc;
    return k;
}

/* Code/main.cpp:733-742: entry n + c is copy c, at its source's site with the
 * source's displacement plus the shift; arrays hold n + ncopies entries */
void oracle_append_copies(long long n, long long ncopies, const int* copy_atom,
                          const double* copy_shift, double* x, double* y, double* z, double* ux,
                          double* uy, double* uz, int* grain)
{
    for (long long c = 0; c < ncopies; c++) {
        const int a = copy_atom[c];
        x[n + c] = x[a]; y[n + c] = y[a]; z[n + c] = z[a];
        if (grain) grain[n + c] = grain[a];
        ux[n + c] = ux[a] + copy_shift[3 * c];
        uy[n + c] = uy[a] + copy_shift[3 * c + 1];
        uz[n + c] = uz[a] + copy_shift[3 * c + 2];
    }
}
