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
 * The slabs are then applied as DISCO's driver applies them: slab entries
 * outside the loop's grain or twin region dropped, vacancy atoms deleted and
 * the rest renumbered in order, interstitial copies appended after the
 * displacement sum with the source's displacement plus a shift -- on the
 * device, as order-keeping compactions, so positions and displacements never
 * go back to the host.
 *
 * Reference interfaces replaced:
 *   disco_hip_matter_type       matterType     (Code/prismatic.cpp:12-18)
 *   disco_hip_prismatic_slab    prismaticSlab  (Code/prismatic.cpp:30-74), with
 *                               loop::contains / insideUV (Code/loop.cpp:281-330)
 *   disco_hip_slab_filter       the grain / twin filter (Code/main.cpp:605-615)
 *   disco_hip_slab_mark_removed the vacancy flags       (Code/main.cpp:626-628)
 *   disco_hip_slab_copies       the copy list of an interstitial slab
 *                               (Code/main.cpp:630-655; the shifts, side * b or
 *                               frankCopyShift's, from the host: disco_geometry.h)
 *   disco_hip_drop_atoms        deletion and renumbering (Code/main.cpp:661-684)
 *   disco_hip_append_copies     the copies after the sum (Code/main.cpp:733-742)
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

/* Keep the slab entries q whose atom a = d_atoms[q] is in the loop's grain and
 * twin region: (d_grain == NULL || d_grain[a] == grain) and (d_twin == NULL ||
 * d_twin[a] == region); entries naming no atom (a outside [0, natoms)) are
 * dropped.  In place, order kept; *count becomes the kept count.  Nothing is
 * done when both label arrays are NULL.  Synchronous on `stream`. */
int disco_hip_slab_filter(disco_ctx *ctx, long long natoms, const int *d_grain, int grain,
                          const int *d_twin, int region, int *d_atoms, int *d_side,
                          long long *count, void *stream);

/* d_removed[d_atoms[q]] = 1 for the count entries of a vacancy slab (entries
 * outside [0, natoms) are ignored).  d_removed: natoms bytes, zeroed by the
 * caller before the first slab.  Asynchronous. */
int disco_hip_slab_mark_removed(disco_ctx *ctx, long long natoms, long long count,
                                const int *d_atoms, unsigned char *d_removed, void *stream);

/* The copies of an interstitial slab: copy q gets source d_atoms[q] and shift
 * shift[0..2] when d_side[q] > 0, shift[3..5] otherwise, written to
 * d_copy_atom[q] and d_copy_shift[3q .. 3q+2] (pass the arrays offset by the
 * copies of earlier loops).  Asynchronous. */
int disco_hip_slab_copies(disco_ctx *ctx, long long count, const int *d_atoms, const int *d_side,
                          const double shift[6], int *d_copy_atom, double *d_copy_shift,
                          void *stream);

/* Delete the flagged atoms: atoms with d_removed[a] == 0 are written in index
 * order to d_xo, d_yo, d_zo (and their label to d_label_o when d_label is not
 * NULL) -- out of place, n entries at most; copies whose source was deleted
 * (or that name no atom) are dropped and the others renumbered, in order, into
 * d_copy_atom_o / d_copy_shift_o.  *n_kept and *ncopies_kept are the new
 * counts.  Synchronous on `stream`. */
int disco_hip_drop_atoms(disco_ctx *ctx, long long n, const unsigned char *d_removed,
                         const double *d_x, const double *d_y, const double *d_z,
                         const int *d_label, double *d_xo, double *d_yo, double *d_zo,
                         int *d_label_o, long long ncopies, const int *d_copy_atom,
                         const double *d_copy_shift, int *d_copy_atom_o, double *d_copy_shift_o,
                         long long *n_kept, long long *ncopies_kept, void *stream);

/* After the displacement sum: entry n + k becomes copy k -- the position (and
 * label, when d_label is not NULL) of its source a = d_copy_atom[k] and the
 * displacement u[a] + shift.  The arrays hold n + ncopies entries; a copy
 * naming no atom (a outside [0, n)) leaves its entry unwritten.
 * Asynchronous. */
int disco_hip_append_copies(disco_ctx *ctx, long long n, long long ncopies,
                            const int *d_copy_atom, const double *d_copy_shift, double *d_x,
                            double *d_y, double *d_z, double *d_ux, double *d_uy, double *d_uz,
                            int *d_label, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* DISCO_PRISMATIC_H */
