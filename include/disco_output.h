/*
 * disco_output.h -- the output stage after the displacement sum (C ABI).
 *
 * DISCO's driver turns (lattice site + displacement) into final positions --
 * wrapped into the periodic box, shifted to [0, L), mapped by the stretch U --
 * and writes them as extended-XYZ and/or LAMMPS data text.  Here both steps
 * run on the device: the final positions are one elementwise pass, and the
 * text of every atom line is formatted by the device (byte for byte the
 * reference's iostream output: fixed, 8 decimals, correctly rounded, ties to
 * even) and streamed to the file through pinned buffers.  Headers are
 * formatted on the host.
 *
 * Reference interfaces replaced:
 *   disco_hip_final_positions  the final-position loop  (Code/main.cpp:797-818)
 *   disco_hip_lammps_box       writeLAMMPS's Gram-Schmidt (Code/main.cpp:64-86)
 *   disco_hip_format_text      the atom lines of the XYZ writer (Code/main.cpp:836-842)
 *                              and of writeLAMMPS (Code/main.cpp:119-127)
 *   disco_hip_write_xyz        the XYZ writer             (Code/main.cpp:820-845)
 *   disco_hip_write_lammps     writeLAMMPS                (Code/main.cpp:50-135)
 *
 * Conventions as in disco_hip.h: device arrays as plain pointers, a stream as
 * void*, int return codes.  Coordinates to be printed must satisfy |v| < 1e10
 * (or be inf / nan, printed as the C library prints them); larger magnitudes
 * are rejected with DISCO_ERR_TOO_LARGE.
 */
#ifndef DISCO_OUTPUT_H
#define DISCO_OUTPUT_H

#include "disco_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define DISCO_TEXT_XYZ 0
#define DISCO_TEXT_LAMMPS 1

/* How atom lines are formatted. */
typedef struct {
    int format;              /* DISCO_TEXT_XYZ or DISCO_TEXT_LAMMPS                        */
    char element[16];        /* XYZ: species column, NUL-terminated, at most 15 chars      */
    const int *d_grain;      /* XYZ: device grain ids, printed + 1; NULL = no column       */
    const int *d_twin;       /* XYZ: device twin ids; NULL = no column                     */
    long long first_id;      /* LAMMPS: id of the first line (atom k has id k + 1)         */
    double frame[9];         /* LAMMPS: rows e1, e2, e3 of disco_hip_lammps_box            */
} disco_text_format;

/* Final positions: r = x + u; each component wrapped once into [-L/2, L/2]
 * (r > L/2: r -= L, else r < -L/2: r += L); then f = U (r + L/2).  U is
 * row-major, NULL = identity.  Device arrays; asynchronous on `stream`. */
int disco_hip_final_positions(disco_ctx *ctx, long long n, const double *d_x, const double *d_y,
                              const double *d_z, const double *d_ux, const double *d_uy,
                              const double *d_uz, const double box[3], const double *U,
                              double *d_fx, double *d_fy, double *d_fz, void *stream);

/* LAMMPS restricted-triclinic box of the deformed box Hp (row-major, columns =
 * box vectors): box6 = (lx, ly, lz, xy, xz, yz), frame9 = rows e1, e2, e3.
 * Host only. */
int disco_hip_lammps_box(const double Hp[9], double box6[6], double frame9[9]);

/* The text of the lines of atoms 0 .. n-1 into device memory d_out (cap
 * bytes); *nbytes = its length.  With d_out == NULL only *nbytes is computed.
 * Synchronous on `stream`.  DISCO_ERR_TOO_LARGE if cap is too small or a
 * coordinate is out of range. */
int disco_hip_format_text(disco_ctx *ctx, const disco_text_format *fmt, long long n,
                          const double *d_fx, const double *d_fy, const double *d_fz,
                          char *d_out, long long cap, long long *nbytes, void *stream);

/* Timing of the text kernels (HIP events on the call's stream): with profiling
 * on, every format call (also inside the file writers) records the durations of
 * its last count, scan and write kernels, read back in ms[0..2]. */
int disco_hip_text_profile(disco_ctx *ctx, int on);
int disco_hip_text_kernel_ms(disco_ctx *ctx, double ms[3]);

/* Whole files from device final positions.  The XYZ header lists Hp's columns
 * as the Lattice string; grain / twin columns are written when given.  The
 * LAMMPS file is the atomic-style data file of writeLAMMPS (tilt line only
 * for a triclinic box).  Synchronous; DISCO_ERR_IO when the file cannot be
 * written. */
int disco_hip_write_xyz(disco_ctx *ctx, const char *path, long long n, const double *d_fx,
                        const double *d_fy, const double *d_fz, const double Hp[9],
                        const char *element, const int *d_grain, const int *d_twin,
                        void *stream);
int disco_hip_write_lammps(disco_ctx *ctx, const char *path, long long n, const double *d_fx,
                           const double *d_fy, const double *d_fz, const double Hp[9],
                           const char *element, double mass, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* DISCO_OUTPUT_H */
