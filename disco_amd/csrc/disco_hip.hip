// disco_hip.hip -- gfx950 kernels and the C ABI of include/disco_hip.h.
//
// Work decomposition (MI355X / CDNA4):
//   * one atom per lane, 256-lane workgroups (4 wave64s), grid = ceil(n / 256);
//     atoms are FP64 SoA in HBM (48 B per atom: x, y, z read once, u written once);
//   * every loop, image and segment of a sample is processed in ONE launch; the
//     loop and segment tables are wave-uniform, so they are read with scalar
//     loads into SGPRs and broadcast to the 64 lanes as instruction operands
//     (no LDS traffic, no per-lane loads of geometry);
//   * FP32C: each (loop, image) pair reduces the atom to q = a - c - delta in
//     FP64, rounds it, and evaluates the segment sum in FP32 (pair_terms.h);
//     the image field is added to an FP64 accumulator.  Lanes inside a loop's
//     near slab (|h| < H, in-plane radius < R + H) take the FP64 evaluation.
//
// Reference behaviour replaced: kroupaGPU and its launch helpers
// (Code/backend_cuda.cu:387-722), the kernels line_kernel / line_kernel_f
// (:302-373) and kroupa_kernel* (:63-291).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#define DISCO_HD __device__ __forceinline__
#include "disco_hip.h"
#include "pair_terms.h"

using namespace disco;

// ------------------------------------------------------------------ device math
struct DevMath {
    DISCO_HD static float rsqrt_f(float x) { return __builtin_amdgcn_rsqf(x); }
    DISCO_HD static float rcp_f(float x) { return __builtin_amdgcn_rcpf(x); }
    DISCO_HD static float log_f(float x) { return logf(x); }
    DISCO_HD static float atan2_f(float y, float x) { return atan2f(y, x); }
    DISCO_HD static float fabs_f(float x) { return fabsf(x); }
    DISCO_HD static float fmax_f(float a, float b) { return fmaxf(a, b); }
    DISCO_HD static int exponent_f(float x) { return __builtin_amdgcn_frexp_expf(x); }
    DISCO_HD static float ldexp_f(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
    DISCO_HD static double sqrt_d(double x) { return sqrt(x); }
    DISCO_HD static double log1p_d(double x) { return log1p(x); }
    DISCO_HD static double atan2_d(double y, double x) { return atan2(y, x); }
    DISCO_HD static double fabs_d(double x) { return fabs(x); }
    DISCO_HD static double fmax_d(double a, double b) { return fmax(a, b); }
    DISCO_HD static int exponent_d(double x) { return __builtin_amdgcn_frexp_exp(x); }
    DISCO_HD static double ldexp_d(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }
};

#undef DISCO_HD
#define DISCO_HD inline
#include "plan_pack.h"

namespace {

constexpr int kBlock = 256;
constexpr int kLooponChunk = 64;   // loopons per FP32 partial sum (flushed to FP64)

// ------------------------------------------------------------------ kernels
// Near predicate of the FP32C mode: the loop image lies within the loop's near
// slab (|h| < hn and in-plane radius^2 < rn2).  Both line kernels evaluate it
// with contraction off, so they take exactly the same decision for every
// (atom, loop, image) and each triple is summed by exactly one of them.
__device__ __forceinline__ bool near_slab(double qx, double qy, double qz, const LineLoop& L)
{
#pragma clang fp contract(off)
    const double h = qx * L.nx + qy * L.ny + qz * L.nz;
    const double r2 = (qx * qx + qy * qy + qz * qz) - h * h;
    return fabs(h) < L.hn && r2 < L.rn2;
}

// FP64 line sum: every (loop, image) in FP64.
__global__ __launch_bounds__(kBlock) void line_fp64_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    double* __restrict__ ux, double* __restrict__ uy, double* __restrict__ uz, int n,
    const LineLoop* __restrict__ loops, int nloops, const SegD* __restrict__ segd, double Lx,
    double Ly, double Lz, int img, int accumulate)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int l = 0; l < nloops; l++) {
        const LineLoop& L = loops[l];
        const double pcx = px - L.cx, pcy = py - L.cy, pcz = pz - L.cz;
        for (int di = -img; di <= img; di++)
            for (int dj = -img; dj <= img; dj++)
                for (int dk = -img; dk <= img; dk++) {
                    double ox = 0.0, oy = 0.0, oz = 0.0;
                    line_image_f64<DevMath>(pcx - di * Lx, pcy - dj * Ly, pcz - dk * Lz,
                                            segd + L.seg0, L.nv, L.bx, L.by, L.bz, ox, oy, oz);
                    sx += ox; sy += oy; sz += oz;
                }
    }
    if (accumulate) { sx += ux[i]; sy += uy[i]; sz += uz[i]; }
    ux[i] = sx; uy[i] = sy; uz[i] = sz;
}

// FP32C far pass: every (loop, image) evaluated in FP32 by every lane (the
// segment loop stays in wave-uniform control flow, so the tables are scalar
// loads); the image field is added in FP64 unless the lane is near.
__global__ __launch_bounds__(kBlock) void line_far_f32_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    double* __restrict__ ux, double* __restrict__ uy, double* __restrict__ uz, int n,
    const LineLoop* __restrict__ loops, int nloops, const SegF* __restrict__ segf, double Lx,
    double Ly, double Lz, int img, int accumulate)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int l = 0; l < nloops; l++) {
        const LineLoop& L = loops[l];
        const double pcx = px - L.cx, pcy = py - L.cy, pcz = pz - L.cz;
        const float bx = (float)L.bx, by = (float)L.by, bz = (float)L.bz;
        for (int di = -img; di <= img; di++)
            for (int dj = -img; dj <= img; dj++)
                for (int dk = -img; dk <= img; dk++) {
                    const double qx = pcx - di * Lx, qy = pcy - dj * Ly, qz = pcz - dk * Lz;
                    float ox = 0.f, oy = 0.f, oz = 0.f;
                    line_image_f32<DevMath>((float)qx, (float)qy, (float)qz, segf + L.seg0,
                                            L.nv, bx, by, bz, ox, oy, oz);
                    if (!near_slab(qx, qy, qz, L)) {
                        sx += (double)ox; sy += (double)oy; sz += (double)oz;
                    }
                }
    }
    if (accumulate) { sx += ux[i]; sy += uy[i]; sz += uz[i]; }
    ux[i] = sx; uy[i] = sy; uz[i] = sz;
}

// FP32C near pass: adds the FP64 field of the near (loop, image) pairs.
__global__ __launch_bounds__(kBlock) void line_near_f64_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    double* __restrict__ ux, double* __restrict__ uy, double* __restrict__ uz, int n,
    const LineLoop* __restrict__ loops, int nloops, const SegD* __restrict__ segd, double Lx,
    double Ly, double Lz, int img)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    bool any = false;
    for (int l = 0; l < nloops; l++) {
        const LineLoop& L = loops[l];
        const double pcx = px - L.cx, pcy = py - L.cy, pcz = pz - L.cz;
        for (int di = -img; di <= img; di++)
            for (int dj = -img; dj <= img; dj++)
                for (int dk = -img; dk <= img; dk++) {
                    const double qx = pcx - di * Lx, qy = pcy - dj * Ly, qz = pcz - dk * Lz;
                    if (near_slab(qx, qy, qz, L)) {
                        double ox = 0.0, oy = 0.0, oz = 0.0;
                        line_image_f64<DevMath>(qx, qy, qz, segd + L.seg0, L.nv, L.bx, L.by,
                                                L.bz, ox, oy, oz);
                        sx += ox; sy += oy; sz += oz;
                        any = true;
                    }
                }
    }
    if (any) { ux[i] += sx; uy[i] += sy; uz[i] += sz; }
}

template <int PREC>
__global__ __launch_bounds__(kBlock) void loopon_sum_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    double* __restrict__ ux, double* __restrict__ uy, double* __restrict__ uz, int n,
    const LooponLoop* __restrict__ loops, int nloops, const float4* __restrict__ pf,
    const double4* __restrict__ pd, double Lx, double Ly, double Lz, int img, int accumulate)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int l = 0; l < nloops; l++) {
        const LooponLoop& L = loops[l];
        const double pcx = px - L.cx, pcy = py - L.cy, pcz = pz - L.cz;
        double s1 = 0.0, s2 = 0.0, s3x = 0.0, s3y = 0.0, s3z = 0.0;
        for (int di = -img; di <= img; di++)
            for (int dj = -img; dj <= img; dj++)
                for (int dk = -img; dk <= img; dk++) {
                    const double qx = pcx - di * Lx, qy = pcy - dj * Ly, qz = pcz - dk * Lz;
                    if (PREC == DISCO_PREC_FP64) {
                        loopon_image_f64<DevMath>(qx, qy, qz, pd + L.p0, L.np, L.mx, L.my, L.mz,
                                                  L.nx, L.ny, L.nz, L.b_n, L.c1, L.c2, s1, s2,
                                                  s3x, s3y, s3z);
                    } else {
                        const float fx = (float)qx, fy = (float)qy, fz = (float)qz;
                        for (int c0 = 0; c0 < L.np; c0 += kLooponChunk) {
                            const int cnt = min(kLooponChunk, L.np - c0);
                            float t1 = 0.f, t2 = 0.f, t3x = 0.f, t3y = 0.f, t3z = 0.f;
                            loopon_image_f32<DevMath>(fx, fy, fz, pf + L.p0 + c0, cnt,
                                                      (float)L.mx, (float)L.my, (float)L.mz,
                                                      (float)L.nx, (float)L.ny, (float)L.nz,
                                                      (float)L.b_n, (float)L.c1, (float)L.c2,
                                                      t1, t2, t3x, t3y, t3z);
                            s1 += t1; s2 += t2; s3x += t3x; s3y += t3y; s3z += t3z;
                        }
                    }
                }
        ax += L.bx * (L.nx * s1 + L.mx * s2 + s3x);
        ay += L.by * (L.ny * s1 + L.my * s2 + s3y);
        az += L.bz * (L.nz * s1 + L.mz * s2 + s3z);
    }
    if (accumulate) { ax += ux[i]; ay += uy[i]; az += uz[i]; }
    ux[i] = ax; uy[i] = ay; uz[i] = az;
}

// u = 0 over [0, n): clears outputs without a memset node (graph-capture safe)
__global__ __launch_bounds__(kBlock) void zero3_kernel(double* __restrict__ a,
                                                       double* __restrict__ b,
                                                       double* __restrict__ c, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) { a[i] = 0.0; b[i] = 0.0; c[i] = 0.0; }
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

// ------------------------------------------------------------------ context
struct disco_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    // cached atom buffers (grow-only)
    double* d_atoms = nullptr;   // x | y | z | ux | uy | uz
    size_t cap_atoms = 0;
};

struct disco_plan {
    disco_ctx* ctx = nullptr;
    int kind = 0;          // 0 line, 1 loopon
    int prec = 0;
    int nloops = 0;        // packed (non-degenerate) loops
    int img = 1;
    double L[3] = {0, 0, 0};
    double work = 0;       // segments (or loopons) per image
    void* d_loops = nullptr;
    void* d_f = nullptr;   // SegF / float4 table
    void* d_d = nullptr;   // SegD / double4 table
    bool profile = false;
    bool profiled_run = false;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
};

namespace {

int fail(disco_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

#define DISCO_HIP_TRY(ctx, call)                                                        \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess)                                                           \
            return fail((ctx), _e == hipErrorOutOfMemory ? DISCO_ERR_OOM : DISCO_ERR_HIP, \
                        std::string(#call) + ": " + hipGetErrorString(_e));             \
    } while (0)

int check_common(disco_ctx* ctx, int n, double Lx, double Ly, double Lz, int img, int prec,
                 int nloops)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (n < 0 || nloops < 0) return fail(ctx, DISCO_ERR_ARG, "negative size");
    if (img < 0) return fail(ctx, DISCO_ERR_ARG, "negative image range");
    if (img > DISCO_MAX_IMG) return fail(ctx, DISCO_ERR_TOO_LARGE, "image range above DISCO_MAX_IMG");
    if (nloops > DISCO_MAX_LOOPS) return fail(ctx, DISCO_ERR_TOO_LARGE, "too many loops");
    if (prec != DISCO_PREC_FP64 && prec != DISCO_PREC_FP32C)
        return fail(ctx, DISCO_ERR_ARG, "unknown precision");
    if (!(std::isfinite(Lx) && std::isfinite(Ly) && std::isfinite(Lz)))
        return fail(ctx, DISCO_ERR_ARG, "non-finite box length");
    return DISCO_OK;
}

int upload_plan(disco_ctx* ctx, disco_plan* p, const void* loops, size_t loop_bytes,
                const void* tf, size_t f_bytes, const void* td, size_t d_bytes)
{
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (loop_bytes) {
        DISCO_HIP_TRY(ctx, hipMalloc(&p->d_loops, loop_bytes));
        DISCO_HIP_TRY(ctx, hipMemcpyAsync(p->d_loops, loops, loop_bytes, hipMemcpyHostToDevice,
                                          ctx->stream));
    }
    if (f_bytes) {
        DISCO_HIP_TRY(ctx, hipMalloc(&p->d_f, f_bytes));
        DISCO_HIP_TRY(ctx, hipMemcpyAsync(p->d_f, tf, f_bytes, hipMemcpyHostToDevice, ctx->stream));
    }
    if (d_bytes) {
        DISCO_HIP_TRY(ctx, hipMalloc(&p->d_d, d_bytes));
        DISCO_HIP_TRY(ctx, hipMemcpyAsync(p->d_d, td, d_bytes, hipMemcpyHostToDevice, ctx->stream));
    }
    DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DISCO_OK;
}

void free_plan(disco_plan* p)
{
    if (!p) return;
    for (hipEvent_t& e : p->ev)
        if (e) (void)hipEventDestroy(e);
    if (p->d_loops) (void)hipFree(p->d_loops);
    if (p->d_f) (void)hipFree(p->d_f);
    if (p->d_d) (void)hipFree(p->d_d);
    delete p;
}

int launch(disco_plan* p, int n, const double* x, const double* y, const double* z, double* ux,
           double* uy, double* uz, int accumulate, hipStream_t s)
{
    disco_ctx* ctx = p->ctx;
    if (n == 0) return DISCO_OK;
    const dim3 grid((unsigned)((n + kBlock - 1) / kBlock)), block(kBlock);
    const bool prof = p->profile && p->nloops > 0;
    p->profiled_run = false;
    if (prof) DISCO_HIP_TRY(ctx, hipEventRecord(p->ev[0], s));
    if (p->nloops == 0) {
        if (!accumulate) hipLaunchKernelGGL(zero3_kernel, grid, block, 0, s, ux, uy, uz, n);
        DISCO_HIP_TRY(ctx, hipGetLastError());
        return DISCO_OK;
    }
    if (p->kind == 0) {
        const LineLoop* lp = (const LineLoop*)p->d_loops;
        const SegF* sf = (const SegF*)p->d_f;
        const SegD* sd = (const SegD*)p->d_d;
        if (p->prec == DISCO_PREC_FP64) {
            hipLaunchKernelGGL(line_fp64_kernel, grid, block, 0, s, x, y, z, ux, uy, uz, n, lp,
                               p->nloops, sd, p->L[0], p->L[1], p->L[2], p->img, accumulate);
        } else {
            hipLaunchKernelGGL(line_far_f32_kernel, grid, block, 0, s, x, y, z, ux, uy, uz, n,
                               lp, p->nloops, sf, p->L[0], p->L[1], p->L[2], p->img, accumulate);
            if (prof) DISCO_HIP_TRY(ctx, hipEventRecord(p->ev[1], s));
            hipLaunchKernelGGL(line_near_f64_kernel, grid, block, 0, s, x, y, z, ux, uy, uz, n,
                               lp, p->nloops, sd, p->L[0], p->L[1], p->L[2], p->img);
        }
    } else {
        const LooponLoop* lp = (const LooponLoop*)p->d_loops;
        const float4* pf = (const float4*)p->d_f;
        const double4* pd = (const double4*)p->d_d;
        if (p->prec == DISCO_PREC_FP64)
            hipLaunchKernelGGL(loopon_sum_kernel<DISCO_PREC_FP64>, grid, block, 0, s, x, y, z,
                               ux, uy, uz, n, lp, p->nloops, pf, pd, p->L[0], p->L[1], p->L[2],
                               p->img, accumulate);
        else
            hipLaunchKernelGGL(loopon_sum_kernel<DISCO_PREC_FP32C>, grid, block, 0, s, x, y, z,
                               ux, uy, uz, n, lp, p->nloops, pf, pd, p->L[0], p->L[1], p->L[2],
                               p->img, accumulate);
    }
    DISCO_HIP_TRY(ctx, hipGetLastError());
    if (prof) {
        if (!(p->kind == 0 && p->prec == DISCO_PREC_FP32C))
            DISCO_HIP_TRY(ctx, hipEventRecord(p->ev[1], s));
        DISCO_HIP_TRY(ctx, hipEventRecord(p->ev[2], s));
        p->profiled_run = true;
    }
    return DISCO_OK;
}

int make_line_plan(disco_ctx* ctx, const disco_line_loop* loops, int nloops, double Lx, double Ly,
                   double Lz, int img, int prec, disco_plan** out, std::vector<double>* work)
{
    LinePack pk;
    const std::string e = pack_line(loops, nloops, pk);
    if (!e.empty()) return fail(ctx, DISCO_ERR_ARG, e);
    disco_plan* p = new (std::nothrow) disco_plan();
    if (!p) return fail(ctx, DISCO_ERR_OOM, "host allocation");
    p->ctx = ctx; p->kind = 0; p->prec = prec; p->img = img;
    p->L[0] = Lx; p->L[1] = Ly; p->L[2] = Lz;
    p->nloops = (int)pk.loops.size();
    p->work = pk.segments;
    const int rc = upload_plan(ctx, p, pk.loops.data(), pk.loops.size() * sizeof(LineLoop),
                               pk.segf.data(), pk.segf.size() * sizeof(SegF), pk.segd.data(),
                               pk.segd.size() * sizeof(SegD));
    if (rc != DISCO_OK) { free_plan(p); return rc; }
    if (work) *work = pk.work;
    *out = p;
    return DISCO_OK;
}

int make_loopon_plan(disco_ctx* ctx, const disco_loopon_loop* loops, int nloops, double Lx,
                     double Ly, double Lz, int img, int prec, disco_plan** out,
                     std::vector<double>* work)
{
    LooponPack pk;
    const std::string e = pack_loopon(loops, nloops, pk);
    if (!e.empty()) return fail(ctx, DISCO_ERR_ARG, e);
    disco_plan* p = new (std::nothrow) disco_plan();
    if (!p) return fail(ctx, DISCO_ERR_OOM, "host allocation");
    p->ctx = ctx; p->kind = 1; p->prec = prec; p->img = img;
    p->L[0] = Lx; p->L[1] = Ly; p->L[2] = Lz;
    p->nloops = (int)pk.loops.size();
    p->work = pk.loopons;
    const int rc = upload_plan(ctx, p, pk.loops.data(), pk.loops.size() * sizeof(LooponLoop),
                               pk.pf.data(), pk.pf.size() * sizeof(float4), pk.pd.data(),
                               pk.pd.size() * sizeof(double4));
    if (rc != DISCO_OK) { free_plan(p); return rc; }
    if (work) *work = pk.work;
    *out = p;
    return DISCO_OK;
}

// Host-buffer sum: upload atoms (and u when accumulating into non-zero u is
// needed: u is always accumulated on the host side), run, download.
int host_sum(disco_ctx* ctx, disco_plan* p, const std::vector<double>& work, int n,
             const double* x, const double* y, const double* z, double* ux, double* uy,
             double* uz, disco_timing* t, int nloops_in)
{
    const double t0 = now_s();
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t need = 6 * (size_t)n;
    if (need > ctx->cap_atoms) {
        if (ctx->d_atoms) DISCO_HIP_TRY(ctx, hipFree(ctx->d_atoms));
        ctx->d_atoms = nullptr;
        ctx->cap_atoms = 0;
        DISCO_HIP_TRY(ctx, hipMalloc(&ctx->d_atoms, need * sizeof(double)));
        ctx->cap_atoms = need;
    }
    double* dx = ctx->d_atoms;
    double *dy = dx + n, *dz = dy + n, *dux = dz + n, *duy = dux + n, *duz = duy + n;
    const size_t bytes = sizeof(double) * (size_t)n;
    DISCO_HIP_TRY(ctx, hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, ctx->stream));
    DISCO_HIP_TRY(ctx, hipMemcpyAsync(dy, y, bytes, hipMemcpyHostToDevice, ctx->stream));
    DISCO_HIP_TRY(ctx, hipMemcpyAsync(dz, z, bytes, hipMemcpyHostToDevice, ctx->stream));
    DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const double t1 = now_s();
    DISCO_HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    int rc = launch(p, n, dx, dy, dz, dux, duy, duz, 0, ctx->stream);
    if (rc != DISCO_OK) return rc;
    DISCO_HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const double t2 = now_s();
    float ms = 0.f;
    DISCO_HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    // download into host staging, then accumulate into the caller's u
    std::vector<double> h(3 * (size_t)n);
    if (n > 0 && p->nloops > 0) {
        DISCO_HIP_TRY(ctx, hipMemcpyAsync(h.data(), dux, 3 * bytes, hipMemcpyDeviceToHost,
                                          ctx->stream));
        DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        for (int a = 0; a < n; a++) {
            ux[a] += h[a];
            uy[a] += h[(size_t)n + a];
            uz[a] += h[2 * (size_t)n + a];
        }
    }
    const double t3 = now_s();
    if (t) {
        t->upload = t1 - t0;
        t->kernels = t2 - t1;
        t->download = t3 - t2;
        t->kernel_ms = ms;
        if (t->loop_sec) {
            double tot = 0.0;
            for (int l = 0; l < nloops_in; l++) tot += work[l];
            for (int l = 0; l < nloops_in; l++)
                t->loop_sec[l] = tot > 0 ? t->kernels * work[l] / tot : 0.0;
        }
    }
    return DISCO_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

const char* disco_hip_strerror(int code)
{
    switch (code) {
        case DISCO_OK: return "success";
        case DISCO_ERR_ARG: return "invalid argument";
        case DISCO_ERR_NO_DEVICE: return "no usable HIP device";
        case DISCO_ERR_HIP: return "HIP runtime error";
        case DISCO_ERR_OOM: return "out of device memory";
        case DISCO_ERR_TOO_LARGE: return "size beyond a documented maximum";
        case DISCO_ERR_STATE: return "call out of order";
        default: return "unknown error code";
    }
}

int disco_hip_device_count(int* count)
{
    if (!count) return DISCO_ERR_ARG;
    *count = 0;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return DISCO_OK;   // no runtime device: 0
    *count = c;
    return DISCO_OK;
}

int disco_hip_device_info(int device, disco_device_info* info)
{
    if (!info) return DISCO_ERR_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || device < 0 || device >= c)
        return DISCO_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return DISCO_ERR_HIP;
    std::memset(info, 0, sizeof(*info));
    std::snprintf(info->name, sizeof(info->name), "%s", prop.name);
    info->compute_units = prop.multiProcessorCount;
    info->global_mem = prop.totalGlobalMem;
    info->lds_per_block = prop.sharedMemPerBlock;
    info->arch_major = prop.major;
    info->arch_minor = prop.minor;
    std::snprintf(info->arch, sizeof(info->arch), "%s", prop.gcnArchName);
    return DISCO_OK;
}

int disco_hip_line_fp64(void) { return 1; }

int disco_hip_open(int device, disco_ctx** out)
{
    if (!out) return DISCO_ERR_ARG;
    *out = nullptr;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c <= 0) return DISCO_ERR_NO_DEVICE;
    if (device < 0 || device >= c) return DISCO_ERR_NO_DEVICE;
    disco_ctx* ctx = new (std::nothrow) disco_ctx();
    if (!ctx) return DISCO_ERR_OOM;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        disco_hip_close(ctx);
        return DISCO_ERR_HIP;
    }
    *out = ctx;
    return DISCO_OK;
}

int disco_hip_close(disco_ctx* ctx)
{
    if (!ctx) return DISCO_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->d_atoms) (void)hipFree(ctx->d_atoms);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return DISCO_OK;
}

const char* disco_hip_last_error(const disco_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

int disco_hip_line_sum(disco_ctx* ctx, int n, const double* x, const double* y, const double* z,
                       double* ux, double* uy, double* uz, const disco_line_loop* loops,
                       int nloops, double Lx, double Ly, double Lz, int img, int precision,
                       disco_timing* t)
{
    int rc = check_common(ctx, n, Lx, Ly, Lz, img, precision, nloops);
    if (rc != DISCO_OK) return rc;
    if (n > 0 && (!x || !y || !z || !ux || !uy || !uz)) return fail(ctx, DISCO_ERR_ARG, "null atom array");
    if (nloops > 0 && !loops) return fail(ctx, DISCO_ERR_ARG, "null loop array");
    disco_plan* p = nullptr;
    std::vector<double> work;
    rc = make_line_plan(ctx, loops, nloops, Lx, Ly, Lz, img, precision, &p, &work);
    if (rc != DISCO_OK) return rc;
    rc = host_sum(ctx, p, work, n, x, y, z, ux, uy, uz, t, nloops);
    free_plan(p);
    return rc;
}

int disco_hip_loopon_sum(disco_ctx* ctx, int n, const double* x, const double* y,
                         const double* z, double* ux, double* uy, double* uz,
                         const disco_loopon_loop* loops, int nloops, double Lx, double Ly,
                         double Lz, int img, int precision, disco_timing* t)
{
    int rc = check_common(ctx, n, Lx, Ly, Lz, img, precision, nloops);
    if (rc != DISCO_OK) return rc;
    if (n > 0 && (!x || !y || !z || !ux || !uy || !uz)) return fail(ctx, DISCO_ERR_ARG, "null atom array");
    if (nloops > 0 && !loops) return fail(ctx, DISCO_ERR_ARG, "null loop array");
    disco_plan* p = nullptr;
    std::vector<double> work;
    rc = make_loopon_plan(ctx, loops, nloops, Lx, Ly, Lz, img, precision, &p, &work);
    if (rc != DISCO_OK) return rc;
    rc = host_sum(ctx, p, work, n, x, y, z, ux, uy, uz, t, nloops);
    free_plan(p);
    return rc;
}

int disco_hip_plan_line(disco_ctx* ctx, const disco_line_loop* loops, int nloops, double Lx,
                        double Ly, double Lz, int img, int precision, disco_plan** plan)
{
    if (!plan) return fail(ctx, DISCO_ERR_ARG, "null plan pointer");
    *plan = nullptr;
    const int rc = check_common(ctx, 0, Lx, Ly, Lz, img, precision, nloops);
    if (rc != DISCO_OK) return rc;
    if (nloops > 0 && !loops) return fail(ctx, DISCO_ERR_ARG, "null loop array");
    return make_line_plan(ctx, loops, nloops, Lx, Ly, Lz, img, precision, plan, nullptr);
}

int disco_hip_plan_loopon(disco_ctx* ctx, const disco_loopon_loop* loops, int nloops, double Lx,
                          double Ly, double Lz, int img, int precision, disco_plan** plan)
{
    if (!plan) return fail(ctx, DISCO_ERR_ARG, "null plan pointer");
    *plan = nullptr;
    const int rc = check_common(ctx, 0, Lx, Ly, Lz, img, precision, nloops);
    if (rc != DISCO_OK) return rc;
    if (nloops > 0 && !loops) return fail(ctx, DISCO_ERR_ARG, "null loop array");
    return make_loopon_plan(ctx, loops, nloops, Lx, Ly, Lz, img, precision, plan, nullptr);
}

int disco_hip_plan_run(disco_plan* plan, int n, const double* d_x, const double* d_y,
                       const double* d_z, double* d_ux, double* d_uy, double* d_uz,
                       int accumulate, void* stream)
{
    if (!plan) return DISCO_ERR_ARG;
    disco_ctx* ctx = plan->ctx;
    if (n < 0) return fail(ctx, DISCO_ERR_ARG, "negative atom count");
    if (n > 0 && (!d_x || !d_y || !d_z || !d_ux || !d_uy || !d_uz))
        return fail(ctx, DISCO_ERR_ARG, "null device array");
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    return launch(plan, n, d_x, d_y, d_z, d_ux, d_uy, d_uz, accumulate ? 1 : 0,
                  (hipStream_t)stream);
}

int disco_hip_plan_profile(disco_plan* plan, int on)
{
    if (!plan) return DISCO_ERR_ARG;
    disco_ctx* ctx = plan->ctx;
    if (on && !plan->ev[0]) {
        DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
        for (hipEvent_t& e : plan->ev) DISCO_HIP_TRY(ctx, hipEventCreate(&e));
    }
    plan->profile = on != 0;
    return DISCO_OK;
}

int disco_hip_plan_kernel_ms(disco_plan* plan, double* ms, int nms)
{
    if (!plan || !ms || nms < 2) return DISCO_ERR_ARG;
    disco_ctx* ctx = plan->ctx;
    if (!plan->profiled_run) return fail(ctx, DISCO_ERR_STATE, "no profiled run yet");
    DISCO_HIP_TRY(ctx, hipEventSynchronize(plan->ev[2]));
    float a = 0.f, b = 0.f;
    DISCO_HIP_TRY(ctx, hipEventElapsedTime(&a, plan->ev[0], plan->ev[1]));
    DISCO_HIP_TRY(ctx, hipEventElapsedTime(&b, plan->ev[1], plan->ev[2]));
    ms[0] = a;
    ms[1] = b;
    return DISCO_OK;
}

double disco_hip_plan_evals_per_atom(const disco_plan* plan)
{
    if (!plan) return 0.0;
    const double im = 2.0 * plan->img + 1.0;
    return plan->work * im * im * im;
}

int disco_hip_plan_free(disco_plan* plan)
{
    free_plan(plan);
    return DISCO_OK;
}

}  // extern "C"
