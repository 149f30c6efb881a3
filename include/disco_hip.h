/*
 * disco_hip.h -- C ABI of the MI355X (gfx950) displacement backend.
 *
 * Plain C types only: host pointers, device pointers as plain pointers, sizes
 * as int, a stream as void* (a hipStream_t, NULL = the null stream).  Every
 * entry point returns an int code (DISCO_OK = 0); disco_hip_strerror() names
 * it and disco_hip_last_error() gives the detail of the last failure on a
 * context.  The C++ host backend (disco_amd/csrc/backend_hip.cpp) maps non-zero
 * codes to the reference's print-and-exit (Code/backend_cuda.cu:45-53).
 *
 * Work and conventions are those of the reference's backends: for every atom,
 * every loop, every periodic image (ii, jj, kk) in [-img, img]^3 and every
 * polyline segment (or loopon), the isotropic displacement is added to the
 * atom's u (Code/backend_cpu.cpp:38-99).  The image sum is "atom at -delta",
 * delta = (ii Lx, jj Ly, kk Lz) (Code/backend_cpu.cpp:41-56); polylines run
 * counter-clockwise about n, V[nv-1] -> V[0] closes them; nu = 0.35.
 */
#ifndef DISCO_HIP_H
#define DISCO_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- codes */
enum {
    DISCO_OK = 0,
    DISCO_ERR_ARG = 1,        /* null pointer, negative size, unknown enum value     */
    DISCO_ERR_NO_DEVICE = 2,  /* no usable HIP device                                */
    DISCO_ERR_HIP = 3,        /* a HIP runtime call failed (detail: last_error)      */
    DISCO_ERR_OOM = 4,        /* device allocation failed                            */
    DISCO_ERR_TOO_LARGE = 5,  /* a size beyond a documented maximum                  */
    DISCO_ERR_STATE = 6,      /* call out of order (e.g. surrogate not trained)      */
    DISCO_ERR_IO = 7          /* a file could not be opened or written               */
};

/* Arithmetic of the pair terms. */
enum {
    DISCO_PREC_FP64 = 0,      /* everything in FP64                                  */
    DISCO_PREC_FP32C = 1      /* FP32 pair terms, FP64 per-image sums, FP64 near the
                                 cut (lanes within the near slab of a loop image)     */
};

/* Limits (checked; DISCO_ERR_TOO_LARGE beyond them). */
#define DISCO_MAX_IMG 8           /* image range +-8 (17^3 images)                    */
#define DISCO_MAX_LOOPS 65536

typedef struct disco_ctx disco_ctx;
typedef struct disco_plan disco_plan;

/* Backend phase times; replaces BackendTiming (Code/backend.hpp:36-39).
 * loop_sec (nloops entries, may be NULL): when given, the host-buffer calls
 * launch each loop on its own and store its measured kernel time (HIP events,
 * seconds), as the reference's kroupaGPU times each loop
 * (Code/backend_cuda.cu:643-700); a degenerate loop (fewer than 3 distinct
 * vertices, no loopons) gets 0.  When NULL, every loop runs in one launch. */
typedef struct {
    double upload;            /* s: allocation + H2D                                  */
    double kernels;           /* s: host wall clock around the kernels (synchronised) */
    double download;          /* s: D2H + free                                        */
    double kernel_ms;         /* ms: HIP-event time of the kernels                    */
    double *loop_sec;
} disco_timing;

/* One loop for the line-integral method; replaces LoopData with line = true
 * (Code/backend.hpp:20-26).  b is the Burgers vector (phys.b * phys.m,
 * Code/backend_cpu.cpp:44-45). */
typedef struct {
    int nv;
    const double *vx, *vy, *vz;
    double c[3];              /* fan apex (loop centre)                               */
    double b[3];
} disco_line_loop;

/* One loop for the loopon method; replaces LoopData with line = false and its
 * LoopPhys (Code/kroupa.hpp:28-36): phys = {bx, by, bz, mx, my, mz, nx, ny, nz, dS}. */
typedef struct {
    int np;
    const double *x, *y, *z;
    double phys[10];
} disco_loopon_loop;

typedef struct {
    char name[256];
    int compute_units;
    size_t global_mem;        /* bytes */
    size_t lds_per_block;     /* bytes */
    int arch_major, arch_minor;
    char arch[64];            /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
} disco_device_info;

/* ---------------------------------------------------------------- device */
/* Number of usable devices; replaces gpuAvailable (Code/backend_cuda.cu:537-541). */
int disco_hip_device_count(int *count);
/* Properties for the banner; replaces printGpuInfo (Code/backend_cuda.cu:543-553). */
int disco_hip_device_info(int device, disco_device_info *info);
/* 1: the FP64 line kernel is built; replaces gpuLineFp64 (Code/backend_cuda.cu:525-532). */
int disco_hip_line_fp64(void);

const char *disco_hip_strerror(int code);

/* ---------------------------------------------------------------- context */
/* A context owns a device, a stream and cached device buffers.  It is
 * thread-compatible: one thread at a time per context. */
int disco_hip_open(int device, disco_ctx **ctx);
int disco_hip_close(disco_ctx *ctx);
const char *disco_hip_last_error(const disco_ctx *ctx);

/* ---------------------------------------------------------------- host buffers
 * One call = the reference's kroupaGPU (Code/backend_cuda.cu:568-722) for the
 * line (resp. loopon) loops: upload the atoms and u, sum every loop in one
 * launch (one launch per loop, each timed, when t->loop_sec is given), add to u
 * on the device, download u.  u is accumulated (u += sum), as kroupaCPU does
 * (Code/backend_cpu.cpp:57-59); with a zeroed u this equals the reference
 * GPU's overwrite (Code/backend_cuda.cu:710-712).  t may be NULL. */
int disco_hip_line_sum(disco_ctx *ctx, int n, const double *x, const double *y,
                       const double *z, double *ux, double *uy, double *uz,
                       const disco_line_loop *loops, int nloops,
                       double Lx, double Ly, double Lz, int img, int precision,
                       disco_timing *t);

int disco_hip_loopon_sum(disco_ctx *ctx, int n, const double *x, const double *y,
                         const double *z, double *ux, double *uy, double *uz,
                         const disco_loopon_loop *loops, int nloops,
                         double Lx, double Ly, double Lz, int img, int precision,
                         disco_timing *t);

/* ---------------------------------------------------------------- device buffers
 * A plan holds one sample's loop geometry on the device (packed segment or
 * loopon tables); running it sums over atoms already resident in HBM.  These
 * split kroupaGPU's per-loop upload + launch (Code/backend_cuda.cu:646-700)
 * into a one-time geometry upload and a launch that reads only the atoms.
 * The device arrays are FP64 SoA; accumulate = 1 adds into u, 0 overwrites.
 * The launch is asynchronous on `stream`. */
int disco_hip_plan_line(disco_ctx *ctx, const disco_line_loop *loops, int nloops,
                        double Lx, double Ly, double Lz, int img, int precision,
                        disco_plan **plan);
int disco_hip_plan_loopon(disco_ctx *ctx, const disco_loopon_loop *loops, int nloops,
                          double Lx, double Ly, double Lz, int img, int precision,
                          disco_plan **plan);
int disco_hip_plan_run(disco_plan *plan, int n, const double *d_x, const double *d_y,
                       const double *d_z, double *d_ux, double *d_uy, double *d_uz,
                       int accumulate, void *stream);
/* (No counterpart in the reference.)  disco_hip_plan_run for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc: the same kernels and results, and no
 * allocation in the call.  The partial sums of a split loopon plan
 * (disco_hip_plan_set_split: 24 B x n x groups), which disco_hip_plan_run keeps
 * in a buffer the plan allocates, go to a static device array of
 * DISCO_NOALLOC_SCRATCH_BYTES that the library holds on each device until the
 * process ends; a larger need fails with DISCO_ERR_TOO_LARGE before anything is
 * launched.  Plans without scratch (line plans, unsplit loopon plans) run as
 * in disco_hip_plan_run.  Every call on one device shares that device's array
 * and the kernels write every partial sum they read, so nothing carries over
 * between calls; calls that overlap on one device (other streams or threads)
 * race and must be serialised by the caller. */
#define DISCO_NOALLOC_SCRATCH_BYTES (64u << 20)
int disco_hip_plan_run_noalloc(disco_plan *plan, int n, const double *d_x, const double *d_y,
                               const double *d_z, double *d_ux, double *d_uy, double *d_uz,
                               int accumulate, void *stream);
/* Per-kernel HIP-event timing of later runs (on = 1; off by default).  A
 * profiled run records events on its stream around each of its kernels. */
int disco_hip_plan_profile(disco_plan *plan, int on);
/* Device times of the last profiled run, waiting for it: ms[0] = the main pass
 * (FP64 sum, FP32 far pass, or loopon sum), ms[1] = the FP64 near pass of the
 * FP32C line mode (0 otherwise).  nms = capacity of ms (>= 2). */
int disco_hip_plan_kernel_ms(disco_plan *plan, double *ms, int nms);
/* Sum only the image shell: images (i, j, k) with max(|i|, |j|, |k|) <= inner
 * are skipped (inner = -1, the default, keeps them all; inner <= img).  A
 * far-shell plan (inner = 1) is what the surrogate replaces and validates. */
int disco_hip_plan_set_inner(disco_plan *plan, int inner);
/* (No counterpart in the reference.)  Cubature form of an FP32C line plan's
 * far pass, for sums over many images: for each loop whose cut a 9-point
 * cubature of the Kroupa term represents (the centre and a ring of eight
 * matching the cut's area, second and fourth radial moments and every moment
 * through degree 5 -- circles and regular polygons of >= 6 sides), a lane at
 * distance r from a loop image's centre takes the exact field for
 * r < k_exact R, the cubature for r > k_rule R and a smooth blend between
 * (R = the cut's radius; the near slab is always exact).  The cubature's error
 * falls as (R/r)^8: for |b| = 2.56 Angstrom about 0.15 (R/r)^8 Angstrom at
 * most.  k_exact <= 0 switches it off (the default); otherwise
 * 2 <= k_exact <= k_rule.  eligible (may be NULL) receives the number of loops
 * that use it; the others stay exact.  Synchronous. */
int disco_hip_plan_set_cubature(disco_plan *plan, double k_exact, double k_rule, int *eligible);
/* (No counterpart in the reference.)  Band form of a cubature plan
 * (disco_hip_plan_set_cubature must be on for it to act): for each loop whose
 * cut a two-ring cubature represents (the centre and two rings of twelve
 * points, a Gauss-Radau rule in r^2 matching the cut's radial moments through
 * I8; used only where it matches every moment of degree <= 9 and the cut's
 * moments of degrees 10 and 11 are isotropic -- circles and regular polygons
 * of twelve sides or more), a lane at distance r >= k_band R of a loop image
 * takes that cubature in place of the exact field (the near slab is always
 * exact; beyond k_exact R the 9-point cubature applies as before).  Its error
 * falls as (R/r)^12: for |b| = 2.56 Angstrom at most ~1.2e-7 Angstrom at
 * r = 3R.  k_band <= 0 switches it off (the default); otherwise
 * k_band >= 2.5.  eligible (may be NULL) receives the number of loops that use
 * it.  Synchronous. */
int disco_hip_plan_set_band(disco_plan *plan, double k_band, int *eligible);
/* Host only: for each loop, {eligible (0/1), centre weight, per-point weight
 * of ring 1, of ring 2, ring radius 1, ring radius 2} of its two-ring cubature
 * (out: 6 doubles per loop). */
int disco_hip_line_band(const disco_line_loop *loops, int nloops, double *out);
/* (No counterpart in the reference.)  Leave images of the +-1 shell out of a
 * line plan's sums: bit 9 (i + 1) + 3 (j + 1) + (k + 1) of skip[l] drops image
 * (i, j, k) of input loop l from every pass (the caller sums it another way,
 * say as a cubature on a grid when it lies far from every atom).  skip = NULL
 * clears every mask.  Synchronous. */
int disco_hip_plan_set_image_skip(disco_plan *plan, const unsigned *skip, int nloops);
/* (No counterpart in the reference.)  The same per octant of the atoms: an
 * atom at (x, y, z) is in octant o = [x >= split[0]] + 2 [y >= split[1]] +
 * 4 [z >= split[2]], and skip[8 l + o] is the mask of input loop l for the
 * atoms of octant o (a grid per octant can then take the images far from that
 * octant's atoms only).  An image left out for an atom must lie beyond the near
 * slab of that atom (as every image far enough for a cubature does).  skip =
 * NULL clears every mask; split may not hold NaN.  Synchronous. */
int disco_hip_plan_set_image_skip_octants(disco_plan *plan, const unsigned *skip, int nloops,
                                          const double *split);
/* Host only: for each loop, {eligible (0/1), centre weight, ring weight, ring
 * radius, cut radius} of its 9-point cubature (out: 5 doubles per loop). */
int disco_hip_line_cubature(const disco_line_loop *loops, int nloops, double *out);
/* (No counterpart.) A loopon plan evaluated at few atoms (say the nodes of a
 * coarse grid) splits its loops into `groups` contiguous groups summed by
 * separate workgroups of one launch, then added in group order (1 = off,
 * <= 256); scratch of 24 B x atoms x groups is kept with the plan. */
int disco_hip_plan_set_split(disco_plan *plan, int groups);
/* Work of one run per atom: segments (or loopons) x images. */
double disco_hip_plan_evals_per_atom(const disco_plan *plan);
int disco_hip_plan_free(disco_plan *plan);

/* ---------------------------------------------------------------- batches
 * Many independent samples -- each with its own atoms, loops and box -- summed
 * by ONE persistent launch per pass, captured in a HIP graph on first run and
 * replayed after.  Workgroups take (sample, atom tile) items from a queue;
 * items are ordered longest first.  Replaces calling kroupaGPU once per sample
 * (Code/backend_cuda.cu:568-722) for batched sample generation. */
typedef struct {
    long long atom_offset;    /* first atom of the sample in the concatenated arrays */
    int n;                    /* its atom count                                       */
    int nloops;
    const disco_line_loop *loops;
    double L[3];              /* its periodic box                                     */
} disco_sample;

/* Launch shape of the persistent kernels. */
typedef struct {
    int block;                /* workgroup size: 64, 128 or 256                      */
    int tile;                 /* atoms per work item, a multiple of block            */
    int groups;               /* persistent workgroups per launch (>= 1)             */
} disco_batch_tuning;

typedef struct disco_batch disco_batch;

int disco_hip_batch_plan(disco_ctx *ctx, const disco_sample *samples, int nsamples, int img,
                         int precision, disco_batch **batch);
int disco_hip_batch_set_tuning(disco_batch *batch, const disco_batch_tuning *tuning);
int disco_hip_batch_get_tuning(const disco_batch *batch, disco_batch_tuning *tuning);
/* Times each candidate launch shape on every `stride`-th sample (stride >= 1)
 * and keeps the fastest; writes it to best (may be NULL) and its time, scaled
 * to the whole batch, to best_ms (may be NULL).  Synchronous; it overwrites u. */
int disco_hip_batch_autotune(disco_batch *batch, const double *d_x, const double *d_y,
                             const double *d_z, double *d_ux, double *d_uy, double *d_uz,
                             int stride, void *stream, disco_batch_tuning *best,
                             double *best_ms);
/* One graph launch over all samples (asynchronous on `stream`). */
int disco_hip_batch_run(disco_batch *batch, const double *d_x, const double *d_y,
                        const double *d_z, double *d_ux, double *d_uy, double *d_uz,
                        int accumulate, void *stream);
/* Image shell of a batch, as disco_hip_plan_set_inner. */
int disco_hip_batch_set_inner(disco_batch *batch, int inner);
/* atom x segment x image evaluations of one run. */
double disco_hip_batch_evals(const disco_batch *batch);
int disco_hip_batch_free(disco_batch *batch);

/* ---------------------------------------------------------------- surrogate
 * The learned far image shell.  The field of a loop over the images with
 * max(|i|,|j|,|k|) in (1, img] is the far shell.  It is modelled as a baseline
 * rule, the loop's cut as a few weighted loopons summed over the far shell by a
 * loopon plan with disco_hip_plan_set_inner (one loopon of the loop's area, or
 * the centre and a ring of six matching the cut's area and radial moments, on
 * a grid with disco_hip_grid_interp), plus a residual that an MLP predicts:
 * 14 features -> 128 -> 128 -> 128 -> 3 (hard swish), layer 1 in FP32, layers
 * 2-4 BF16 on MFMA with FP32 accumulation, output scaled per loop by
 * disco_hip_surrogate_scale.  It is trained (Adam, FP32 masters) by the
 * library's own forward and backward kernels.  The host pipeline accepts the
 * surrogate only when its error on a validation sample of atoms against the
 * exact far shell is within the tolerance, else sums the exact far shell; the
 * validation is statistical (sampled atoms).  No counterpart in the
 * reference, which sums +-1 images only (Code/backend_cpu.cpp:51-53). */
#define DISCO_SURR_IN 14
#define DISCO_SURR_OUT 3

typedef struct {
    double c[3];              /* loop centre                                          */
    double n[3];              /* unit plane normal                                    */
    double b[3];              /* Burgers vector                                       */
    double radius;            /* radius (circumradius of a polygon), Angstrom         */
    int shape;                /* LoopShape (Code/loop.hpp:13): 0 circle, 1 rectangle,
                                 2 parallelogram, 3 polygon                           */
} disco_surrogate_loop;

typedef struct disco_surrogate disco_surrogate;

int disco_hip_surrogate_num_params(void);
/* Features of atoms p (n x 3, host) for one loop in box L (3): n x 14 floats. */
int disco_hip_surrogate_features(const double *p, int n, const disco_surrogate_loop *loop,
                                 const double *L, float *out);
/* Output scale of a loop: |b| (radius / mean(L))^4. */
double disco_hip_surrogate_scale(const disco_surrogate_loop *loop, const double *L);
int disco_hip_surrogate_create(disco_ctx *ctx, unsigned seed, disco_surrogate **sur);
int disco_hip_surrogate_free(disco_surrogate *sur);
/* FP32 master parameters (disco_hip_surrogate_num_params floats). */
int disco_hip_surrogate_get_params(disco_surrogate *sur, float *out);
int disco_hip_surrogate_set_params(disco_surrogate *sur, const float *in);
/* Output gain: the surrogate's output is gain * network (default 1); training
 * on targets divided by their RMS and setting gain = that RMS keeps the
 * network's outputs of order one. */
int disco_hip_surrogate_set_gain(disco_surrogate *sur, double gain);
double disco_hip_surrogate_get_gain(const disco_surrogate *sur);
/* gain * network output for n host feature rows (n x 14) into out (n x 3). */
int disco_hip_surrogate_forward(disco_surrogate *sur, const float *feats, int n, float *out);
/* Mean squared error of the network (without the gain) over n rows and its
 * gradient (all parameters). */
int disco_hip_surrogate_loss_grad(disco_surrogate *sur, const float *feats,
                                  const float *targets, int n, float *loss, float *grads);
/* Adam over a host data set (features n x 14, targets n x 3), cosine-decayed
 * learning rate from lr to lr / 100; final_loss = last batch's loss. */
int disco_hip_surrogate_train(disco_surrogate *sur, const float *feats, const float *targets,
                              int n, int epochs, int batch, float lr, unsigned seed,
                              float *final_loss);
/* u += sum over loops of gain * scale * network(features) for device atoms.  The
 * whole network runs in one fused kernel (weights in LDS, activations in
 * registers); all atoms in one launch (chunk_rows <= 0) or chunk_rows / nloops
 * atoms per launch.  Synchronous on `stream`. */
int disco_hip_surrogate_apply(disco_surrogate *sur, int natoms, const double *d_x,
                              const double *d_y, const double *d_z, double *d_ux, double *d_uy,
                              double *d_uz, const disco_surrogate_loop *loops, int nloops,
                              const double *L, int chunk_rows, void *stream);

/* ---------------------------------------------------------------- lattice
 * Perfect cubic lattice written straight into device memory, in the reference's
 * atom order and bit for bit its coordinates: replaces generateLattice
 * (Code/lattice.hpp:35-51; called at Code/main.cpp:576-584) and the host-to-
 * device copy of the positions.  lattice: DISCO_LATTICE_BCC / _FCC; the
 * caller allocates disco_hip_lattice_atoms() doubles per array; 0 <= nx, ny,
 * nz <= 65535.  Asynchronous on `stream`. */
#define DISCO_LATTICE_BCC 0
#define DISCO_LATTICE_FCC 1
long long disco_hip_lattice_atoms(int lattice, int nx, int ny, int nz);
int disco_hip_lattice(disco_ctx *ctx, int lattice, int nx, int ny, int nz, double a0,
                      double *d_x, double *d_y, double *d_z, void *stream);

/* ---------------------------------------------------------------- grid field
 * (No counterpart in the reference.)  u += (accumulate = 1) or u = (0) a field
 * tabulated on a uniform grid, interpolated at n device atoms by the tensor
 * product of `order`-node Lagrange interpolants: order 4 (cubic, exact for
 * polynomials of degree <= 3 along each axis) or 6 (quintic, degree <= 5).
 * g_x, g_y, g_z: nx * ny * nz doubles each, x fastest; node (i, j, k) sits at
 * origin + (i, j, k) h.  nx, ny, nz >= order; an atom outside the nodes' span
 * uses the nearest stencil.  Used for the smooth far images of a wide image
 * sum.  Asynchronous on `stream`. */
int disco_hip_grid_interp(disco_ctx *ctx, int n, const double *d_x, const double *d_y,
                          const double *d_z, double *d_ux, double *d_uy, double *d_uz,
                          const double *d_gx, const double *d_gy, const double *d_gz, int nx,
                          int ny, int nz, const double *origin, double h, int order,
                          int accumulate, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* DISCO_HIP_H */
