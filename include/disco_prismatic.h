/*
 * disco_prismatic.h -- prismatic slab selection on the device (C ABI).
 *
 * A loop whose Burgers vector has a component along its normal (b.n != 0)
 * removes (vacancy) or adds (interstitial) the atoms of a slab: those within
 * k = |b.n| / d atomic planes of the cut and inside the contour.  DISCO's
 * driver finds them with an O(atoms) host scan per loop; here the predicate --
 * minimum image, signed distance to the cut, the loop's in-plane contour test
 * -- runs on the device for every atom, with the reference's rounding (the
 * products it fuses are fused here, divisions and square roots are correctly
 * rounded), and the selected atoms are compacted in ascending index order.
 * The result is the reference's slab bit for bit: same atoms, same order,
 * same sides.
 *
 * Reference interfaces replaced:
 *   disco_hip_matter_type     matterType     (Code/prismatic.cpp:12-18)
 *   disco_hip_prismatic_slab  prismaticSlab  (Code/prismatic.cpp:30-74), with
 *                             loop::contains / insideUV (Code/loop.cpp:281-330)
 *
 * The loop is passed as its constructed state (disco_loop_state of
 * disco_geometry.h, e.g. from disco_geo_loop_state).
 */
#ifndef DISCO_PRISMATIC_H
#define DISCO_PRISMATIC_H

#include "disco_geometry.h"
#include "disco_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DISCO_MATTER_NONE 0          /* glide loop (|m.n| < 1e-6): no slab        */
#define DISCO_MATTER_VACANCY 1       /* m.n > 0: the slab atoms are removed       */
#define DISCO_MATTER_INTERSTITIAL 2  /* m.n < 0: the slab atoms are duplicated    */

typedef struct {
    int type;            /* DISCO_MATTER_*                                          */
    int planes;          /* k: atomic planes in the slab                            */
    double height;       /* |b.n_hat| (Angstrom)                                    */
    double spacing;      /* d (Angstrom)                                            */
    double b[3];         /* Burgers vector |b| m_hat                                */
    long long count;     /* slab atoms                                              */
} disco_slab_info;

/* matterType of a constructed loop (host only). */
int disco_hip_matter_type(const disco_loop_state *loop);

/* The slab of one loop among n device atoms (x, y, z in a periodic box
 * (lx, ly, lz) centred on the origin; spacing = d of the loop's plane family,
 * from snapToMidPlane).  Writes the atom indices (ascending) to d_atoms and
 * the side of each (+1 / -1: sign of the distance to the cut) to d_side, at
 * most cap entries; info->count is the slab size.  Synchronous on `stream`.
 * DISCO_ERR_ARG when |b.n| is not a whole number of spacings (where the
 * reference aborts); DISCO_ERR_TOO_LARGE when count > cap (info filled). */
int disco_hip_prismatic_slab(disco_ctx *ctx, const disco_loop_state *loop, double spacing,
                             long long n, const double *d_x, const double *d_y,
                             const double *d_z, double lx, double ly, double lz, int *d_atoms,
                             int *d_side, long long cap, disco_slab_info *info, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* DISCO_PRISMATIC_H */
