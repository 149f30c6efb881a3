/* disco_geometry.h -- host producer of the backend's loop inputs (C ABI).
 *
 * The loop geometry that DISCO's driver builds before every sum -- closed
 * polylines, loopon grids, the mid-plane snap of loop centres -- restated in
 * C++ for batched sample generation (many samples, thousands of loops): the
 * same formulas and operation order as the reference's loop code, with the
 * products fused where the reference's host build fuses them, so the results
 * are bit-identical to the reference's (tested against the compiled
 * reference's outputs).  Host code only; no device is needed.
 *
 * Reference interfaces replaced:
 *   disco_geo_regular_polygon  regularPolygon      (Code/loop.cpp:585-608)
 *   disco_geo_snap_mid_plane   snapToMidPlane      (Code/loop.cpp:342-374)
 *   disco_geo_polyline         loop::polyline      (Code/loop.cpp:478-580)
 *   disco_geo_loopons          loop::generateLoopons / insideUV
 *                              (Code/loop.cpp:253-316)
 *   disco_geo_phys             the LoopPhys of a loop (Code/loop.cpp:25-29)
 *   disco_geo_loop_state       the loop constructors    (Code/loop.cpp:13-244)
 *   disco_geo_polylines        the driver's per-loop LoopData assembly
 *                              (Code/main.cpp:689-711), for many loops at once
 *   disco_geo_classify_burgers classifyBurgers     (Code/lattice.hpp:189-212)
 *   disco_geo_frank_copy_shift frankCopyShift      (Code/lattice.hpp:239-256)
 *   disco_geo_copy_shifts      the copy shifts of an interstitial loop
 *                              (Code/main.cpp:633-650)
 */
#ifndef DISCO_GEOMETRY_H
#define DISCO_GEOMETRY_H

#ifdef __cplusplus
extern "C" {
#endif

#define DISCO_SHAPE_CIRCLE 0
#define DISCO_SHAPE_POLYGON 1        /* vertices given as offsets from the centre */
#define DISCO_SHAPE_RECT 2
#define DISCO_SHAPE_PARALLELOGRAM 3

#define DISCO_LATTICE_BCC_GEO 0
#define DISCO_LATTICE_FCC_GEO 1

typedef struct {
    int shape;                 /* DISCO_SHAPE_*                                        */
    double center[3];
    double m[3];               /* Burgers direction in units of a0/2: |b| = a0 |m| / 2 */
    double n[3];               /* plane normal                                         */
    double a0;
    double radius;             /* circle                                               */
    double half_m, half_p;     /* rect, parallelogram: half extents along the axes     */
    double dir1[3], dir2[3];   /* parallelogram edge directions                        */
    int nvert;                 /* polygon                                              */
    const double *vert;        /* polygon: 3 * nvert offsets from the centre           */
    double corner_r;           /* rect, polygon: corner rounding radius (0 = sharp)    */
    double axis[3];            /* rect: Side (the half_m axis); polygon: grid axis u;  */
                               /* all zero = the default axis                          */
} disco_loop_shape;

/* A loop as the reference's constructors leave it (class loop's state,
 * Code/loop.hpp:18-45): what the prismatic slab selection reads. */
#define DISCO_MAX_STATE_VERT 128
typedef struct {
    int shape;                 /* DISCO_SHAPE_*                                        */
    int nvert;                 /* polygon vertices                                     */
    double center[3];
    double m[3], n[3];         /* unit Burgers direction and plane normal              */
    double u[3], v[3];         /* in-plane grid axes (oblique for a parallelogram)     */
    double bmag;               /* |b|                                                  */
    double radius, half_m, half_p, r_corner;
    double pu[DISCO_MAX_STATE_VERT], pv[DISCO_MAX_STATE_VERT];   /* polygon P (u, v)   */
    double qu[DISCO_MAX_STATE_VERT], qv[DISCO_MAX_STATE_VERT];   /* inner polygon Q    */
    double nu[DISCO_MAX_STATE_VERT], nv[DISCO_MAX_STATE_VERT];   /* inward side normals */
} disco_loop_state;

/* Return values: >= 0 the number of points (written when <= cap), < 0 an
 * error: DISCO_GEO_ERR_ARG for an invalid shape or argument. */
#define DISCO_GEO_ERR_ARG (-1)

/* Vertices (3 x sides doubles, xyz interleaved) of a regular polygon of
 * circumradius R in the plane of n, as offsets from its centre. */
int disco_geo_regular_polygon(int sides, double R, const double m[3], const double n[3],
                              double *out);
/* Centre moved along n onto the nearest mid-plane between (hkl) planes of the
 * box-centred lattice (nx, ny, nz cells); n integer Miller indices. */
int disco_geo_snap_mid_plane(const double c[3], const double n[3], double a0, int nx, int ny,
                             int nz, int lattice, double out[3], double *spacing);
/* Closed contour, counter-clockwise about n: a circle as nseg chords of the
 * area-preserving radius (nseg <= 0: as many as the sagitta eps needs), a
 * sharp shape by its corners and a rounded corner by chords of sagitta <= eps,
 * or (subdivide > 0, sharp shapes) each side cut into collinear pieces so the
 * contour has `subdivide` segments. */
int disco_geo_polyline(const disco_loop_shape *s, double eps, int nseg, int subdivide,
                       double *vx, double *vy, double *vz, int cap);
/* Loopon grid of spacing da inside the loop. */
int disco_geo_loopons(const disco_loop_shape *s, double da, double *x, double *y, double *z,
                      int cap);
/* phys = (|b|, |b|, |b|, m/|m|, n/|n|, dS) with dS = da^2 (times |dir1 x dir2| for a
 * parallelogram), or 0 for da <= 0. */
int disco_geo_phys(const disco_loop_shape *s, double da, double phys[10]);
/* The loop's state after construction (initCommon, defaultAxes, the shape's
 * constructor: Code/loop.cpp:13-244).  Returns 1, or DISCO_GEO_ERR_ARG for an
 * invalid shape (also a polygon that is not strictly convex or has more than
 * DISCO_MAX_STATE_VERT vertices). */
int disco_geo_loop_state(const disco_loop_shape *s, disco_loop_state *out);
/* Many loops' polylines into concatenated arrays in one call:
 * offsets[l] .. offsets[l + 1] are loop l's vertices; nseg / subdivide per
 * loop (may be NULL: 0).  Returns the total vertex count (arrays written when
 * it is <= cap; offsets always). */
long long disco_geo_polylines(const disco_loop_shape *shapes, int nloops, double eps,
                              const int *nseg, const int *subdivide, long long *offsets,
                              double *vx, double *vy, double *vz, long long cap);

/* Kind of the Burgers vector b = (a0/2) m of a loop of normal n (crystal
 * frame, Miller indices): a lattice translation, an FCC Frank (2/3<111>, n
 * along m) or Shockley (1/3<112> in a {111} plane) partial, or neither. */
#define DISCO_BURGERS_PERFECT 0
#define DISCO_BURGERS_FRANK 1
#define DISCO_BURGERS_SHOCKLEY 2
#define DISCO_BURGERS_OTHER 3
int disco_geo_classify_burgers(int lattice, const double m[3], const double n[3]);
/* Shift of the copy of an interstitial Frank-loop atom from its source (FCC,
 * crystal frame): v = side * b plus the in-plane Shockley vector a0/6<112>
 * (normal to nhat) for which the sum is not a lattice translation, so the
 * inserted {111} layer takes the third stacking position.  out = v when no
 * such vector exists (nhat not a <111> direction). */
void disco_geo_frank_copy_shift(const double v[3], const double nhat[3], double a0,
                                double out[3]);
/* The two copy shifts of an interstitial loop (lab frame): shift[0..2] for the
 * atoms on the +n side of the cut, shift[3..5] for the -n side.  v = +b / -b
 * (b = |b| m, lab frame); for a Frank loop (frank != 0) frankCopyShift of v in
 * the crystal frame, with the loop's unit normal n, brought back to the lab.
 * R: the grain's orientation, rows = the crystal directions of the box axes
 * (a crystal vector c is R c in the lab), row-major; NULL for an unrotated
 * crystal. */
void disco_geo_copy_shifts(const double b[3], int frank, const double n[3], const double *R,
                           double a0, double shift[6]);

#ifdef __cplusplus
}
#endif
#endif /* DISCO_GEOMETRY_H */
