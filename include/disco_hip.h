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
    DISCO_ERR_STATE = 6       /* call out of order (e.g. surrogate not trained)      */
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
 * loop_sec (nloops entries, may be NULL) receives the kernel time apportioned
 * to each loop by its share of the work, since all loops run in one launch. */
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
 * line (resp. loopon) loops: upload the atoms, sum every loop in one launch,
 * download.  u is accumulated (u += sum), as kroupaCPU does
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
/* Per-kernel HIP-event timing of later runs (on = 1; off by default).  A
 * profiled run records events on its stream around each of its kernels. */
int disco_hip_plan_profile(disco_plan *plan, int on);
/* Device times of the last profiled run, waiting for it: ms[0] = the main pass
 * (FP64 sum, FP32 far pass, or loopon sum), ms[1] = the FP64 near pass of the
 * FP32C line mode (0 otherwise).  nms = capacity of ms (>= 2). */
int disco_hip_plan_kernel_ms(disco_plan *plan, double *ms, int nms);
/* Work of one run per atom: segments (or loopons) x images. */
double disco_hip_plan_evals_per_atom(const disco_plan *plan);
int disco_hip_plan_free(disco_plan *plan);

#ifdef __cplusplus
}
#endif
#endif /* DISCO_HIP_H */
