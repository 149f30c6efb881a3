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

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#define DISCO_HD __device__ __forceinline__
#include "ctx_internal.h"
#include "disco_hip.h"
#include "pair_terms.h"

using namespace disco;

// ------------------------------------------------------------------ device math
struct DevMath {
    DISCO_HD static float rsqrt_f(float x) { return __builtin_amdgcn_rsqf(x); }
    DISCO_HD static float rcp_f(float x) { return __builtin_amdgcn_rcpf(x); }
    // natural log from the hardware log2 (x >= 5/3 here: no denormal inputs)
    DISCO_HD static float log_f(float x) { return __builtin_amdgcn_logf(x) * 0.693147181f; }
    DISCO_HD static float atan2_f(float y, float x) { return atan2f(y, x); }
    DISCO_HD static float fabs_f(float x) { return fabsf(x); }
    DISCO_HD static float fmax_f(float a, float b) { return fmaxf(a, b); }
    DISCO_HD static float fmin_f(float a, float b) { return fminf(a, b); }
    DISCO_HD static float copysign_f(float a, float b) { return copysignf(a, b); }
    DISCO_HD static int exponent_f(float x) { return __builtin_amdgcn_frexp_expf(x); }
    DISCO_HD static float ldexp_f(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
    DISCO_HD static double sqrt_d(double x) { return sqrt(x); }
    // 1/sqrt(x) and 1/x: the hardware estimates refined by two Newton steps
    // (quadratic convergence from ~2^-22 to rounding level)
    DISCO_HD static double rsqrt_d(double x)
    {
        const double h = 0.5 * x;
        double y = __builtin_amdgcn_rsq(x);
        y = fma(y, fma(-h * y, y, 0.5), y);
        y = fma(y, fma(-h * y, y, 0.5), y);
        return y;
    }
    DISCO_HD static double rcp_d(double x)
    {
        double y = __builtin_amdgcn_rcp(x);
        y = fma(y, fma(-x, y, 1.0), y);
        y = fma(y, fma(-x, y, 1.0), y);
        return y;
    }
    // the same with one Newton step, for the per-segment terms of the FP64
    // walks: measured on gfx950 the hardware estimates err by at most 5.2e-8
    // (rsq) and 4.6e-8 (rcp) relative, one step leaves 4.1e-15 and 2.2e-15
    DISCO_HD static double rsqrt_d1(double x)
    {
        const double y = __builtin_amdgcn_rsq(x);
        return fma(y, fma(-0.5 * x * y, y, 0.5), y);
    }
    DISCO_HD static double rcp_d1(double x)
    {
        const double y = __builtin_amdgcn_rcp(x);
        return fma(y, fma(-x, y, 1.0), y);
    }
    DISCO_HD static double log1p_d(double x) { return log1p(x); }
    DISCO_HD static double atan2_d(double y, double x) { return atan2(y, x); }
    DISCO_HD static double fabs_d(double x) { return fabs(x); }
    DISCO_HD static bool signbit_d(double x) { return signbit(x); }
    DISCO_HD static double fmax_d(double a, double b) { return fmax(a, b); }
    DISCO_HD static int exponent_d(double x) { return __builtin_amdgcn_frexp_exp(x); }
    DISCO_HD static double ldexp_d(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }
    // wave-uniform votes over the active lanes
    DISCO_HD static bool any(bool p) { return __any(p); }
    DISCO_HD static bool all(bool p) { return __all(p); }
};

#undef DISCO_HD
#define DISCO_HD inline
#include "plan_pack.h"

namespace {

constexpr int kBlock = 256;
constexpr int kLooponChunk = 64;   // loopons per FP32 partial sum (flushed to FP64)
constexpr double kMaxShiftBytes = 1073741824.0;   // a plan's image shift table

// ------------------------------------------------------------------ per-atom sums
// The lane in the frame (e1, e2, n) of loop L: Q = R (p - c), and an image's
// offset Q - ((di X + dj Y) + dk Z) with the box edges X, Y, Z of the loop
// record; the shift (di X + dj Y) + dk Z is wave-uniform and comes from the
// plan's table in the FP64 segment table (pack_shifts: slot L.sh0 on, read with
// scalar loads).  Every line pass takes it from there and Q with contraction
// off, so the near predicate (|h| < hn and in-plane radius^2 < rn2) takes
// exactly the same decision in each and every (atom, loop, image) is summed by
// exactly one of them.
struct FrameAtom {
    double Q1, Q2, Q3;
};

__device__ __forceinline__ void frame_atom(double pcx, double pcy, double pcz, const LineLoop& L,
                                           FrameAtom& f)
{
#pragma clang fp contract(off)
    f.Q1 = (pcx * L.e1x + pcy * L.e1y) + pcz * L.e1z;
    f.Q2 = (pcx * L.e2x + pcy * L.e2y) + pcz * L.e2z;
    f.Q3 = (pcx * L.nx + pcy * L.ny) + pcz * L.nz;
}

__device__ __forceinline__ bool frame_image(const FrameAtom& f, int di, int dj, int dk, int img,
                                            const LineLoop& L, const SegD* __restrict__ segd,
                                            double& q1, double& q2, double& q3)
{
#pragma clang fp contract(off)
    const int n = 2 * img + 1;
    const double* D = reinterpret_cast<const double*>(segd + L.sh0) +
                      4 * (((di + img) * n + (dj + img)) * n + (dk + img));
    q1 = f.Q1 - D[0];
    q2 = f.Q2 - D[1];
    q3 = f.Q3 - D[2];
    return fabs(q3) < L.hn && q1 * q1 + q2 * q2 < L.rn2;
}

// FP64 field of one image of a loop flat to double resolution, evaluated in
// the loop frame and added in global axes: b fo + o1 e1 + o2 e2 + o3 n.
__device__ __forceinline__ void add_flat_f64(double q1, double q2, double q3, const LineLoop& L,
                                             const SegD* __restrict__ segd, bool skip,
                                             double& sx, double& sy, double& sz)
{
    double fo, o1, o2, o3;
    const SegPD* sp = reinterpret_cast<const SegPD*>(segd + L.seg0);
    if (L.kind == LOOP_GLIDE)
        line_image_flat_f64<DevMath, true>(q1, q2, q3, sp, L.nv, fo, o1, o2, o3);
    else
        line_image_flat_f64<DevMath, false>(q1, q2, q3, sp, L.nv, fo, o1, o2, o3);
    if (skip) return;                              // a lane that leaves the image out
    sx += L.bx * fo + ((o1 * L.e1x + o2 * L.e2x) + o3 * L.nx);
    sy += L.by * fo + ((o1 * L.e1y + o2 * L.e2y) + o3 * L.ny);
    sz += L.bz * fo + ((o1 * L.e1z + o2 * L.e2z) + o3 * L.nz);
}

// Range of image indices along one axis that can hold a near image of the
// loop for this lane: a near image has |q| < sqrt(rn2 + hn^2) < reach, so its
// offset d L lies within reach of the apex offset pc along every axis.
__device__ __forceinline__ void near_range(double pc, double iL, double reach, int img, int& lo,
                                           int& hi)
{
    if (iL == 0.0) { lo = -img; hi = img; return; }
    const double a = fmin(fmax(ceil((pc - reach) * iL), -img - 1.0), img + 1.0);
    const double b = fmin(fmax(floor((pc + reach) * iL), -img - 1.0), img + 1.0);
    lo = max((int)a, -img);
    hi = min((int)b, img);
}

// The three line passes over loops [l0, l1) and images [-img, img]^3 (kk
// fastest) for the atom at p; loop, segment and image data are wave-uniform.
enum { PASS_FP64 = 0, PASS_FAR = 1, PASS_NEAR = 2 };
// PASS_FP64 loop filter (the KIND template argument of that pass): only the
// loops in global axes (Seg<double>), or only those flat to double resolution
// (SegPD, the loop-frame walk); each kernel compiles one path, and so needs
// fewer registers.
enum { FP64_GLOBAL = 1, FP64_FRAME = 2 };

// Images a line pass leaves out.  The inner shell (max(|i|,|j|,|k|) <= inner)
// is left out for every lane; an image of [-1, 1]^3 is left out for the lanes
// whose skip mask holds its bit (skip_bit): the mask of the lane's octant
// (lane_skip).  A wave passes over an image only when every lane leaves it out;
// otherwise the lanes that leave it out evaluate it with the others and drop
// the result.
__device__ __forceinline__ bool inner_out(int di, int dj, int dk, int inner)
{
    return max(abs(di), max(abs(dj), abs(dk))) <= inner;
}

__device__ __forceinline__ bool skip_bit(unsigned mask, int di, int dj, int dk)
{
    return max(abs(di), max(abs(dj), abs(dk))) <= 1 &&
           ((mask >> ((di + 1) * 9 + (dj + 1) * 3 + (dk + 1))) & 1u);
}

// Whether the lane leaves image (di, dj, dk) of loop L out (lm: its lane_skip
// mask); a wave-uniform branch keeps loops that leave nothing out free of it.
__device__ __forceinline__ bool lane_skips_image(const LineLoop& L, unsigned lm, int di, int dj,
                                                 int dk)
{
    return L.skip_any ? skip_bit(lm, di, dj, dk) : false;
}

// The octant of the atom for the per-octant skip masks (all loops of a plan
// hold the same planes osx, osy, osz).
__device__ __forceinline__ int atom_octant(double px, double py, double pz, const LineLoop& L)
{
    return (px >= L.osx ? 1 : 0) | (py >= L.osy ? 2 : 0) | (pz >= L.osz ? 4 : 0);
}

// The lane's skip mask of loop L: oct_skip[oct] selected from the eight
// wave-uniform masks (scalar loads; no per-lane memory access); 0 when the
// loop leaves nothing out.
__device__ __forceinline__ unsigned lane_skip(const LineLoop& L, int oct)
{
    if (!L.skip_any) return 0u;
    const unsigned a = (oct & 1) ? L.oct_skip[1] : L.oct_skip[0];
    const unsigned b = (oct & 1) ? L.oct_skip[3] : L.oct_skip[2];
    const unsigned c = (oct & 1) ? L.oct_skip[5] : L.oct_skip[4];
    const unsigned d = (oct & 1) ? L.oct_skip[7] : L.oct_skip[6];
    const unsigned e = (oct & 2) ? b : a, g = (oct & 2) ? d : c;
    return (oct & 4) ? g : e;
}

// FP32 far pass of one flat loop: every lane evaluates every image (uniform
// control flow keeps the segment table in scalar loads); near lanes discard
// it.  Each image's field is turned to global axes in float and added to the
// FP64 sums.  CUBATURE: for a loop whose 9-point cubature matches its cut
// (rule_ok), each lane weighs the exact field by 1 - w and the cubature's by w,
// w = rule_window(|q|^2) (0 near the loop, 1 far from it); a wave evaluates the
// exact field only when one of its lanes (outside the near slab) has w < 1,
// and the cubature only when one has w > 0.  With the band form on
// (disco_hip_plan_set_band), the two-ring cubature takes the exact field's
// place for lanes with |q|^2 >= rtb.
// One image (di, dj, dk) of a flat loop in the FP32 far pass (SKIPS: the plan
// leaves images out for some atoms).
template <int KIND, bool CUBATURE, bool SKIPS>
__device__ __forceinline__ void far_flat_image(const FrameAtom& f, const LineLoop& L,
                                               const SegP* __restrict__ s,
                                               const SegD* __restrict__ segd, int img, int di,
                                               int dj, int dk, unsigned lm, double& sx,
                                               double& sy, double& sz)
{
    const bool skip = SKIPS ? lane_skips_image(L, lm, di, dj, dk) : false;
    if (SKIPS && L.skip_any && DevMath::all(skip)) return;
    double q1, q2, q3;
    // a lane that leaves the image out counts as near: the far pass
    // drops its result (the near pass never sees such an image)
    const bool near = frame_image(f, di, dj, dk, img, L, segd, q1, q2, q3) || skip;
    const float g1 = (float)q1, g2 = (float)q2, g3 = (float)q3;
    float fo = 0.f, o1 = 0.f, o2 = 0.f, o3 = 0.f;
    if (!CUBATURE || !L.rule_ok) {
        line_image_flat_f32<DevMath, KIND>(g1, g2, g3, s, L.nv, fo, o1, o2, o3);
    } else {
        const float t = g1 * g1 + (g2 * g2 + g3 * g3);
        // a hard switch at rt0 (k_rule = k_exact, rwinv huge): the
        // window is the step that rule_window gives there, and no
        // lane blends
        const bool hard = L.rwinv >= 1e29f;
        const float w = hard ? (t > L.rt0 ? 1.0f : 0.0f) : rule_window(t, L);
        // the two-ring rule stands in for the exact field beyond rtb
        // (never inside the near slab; rtb is huge when it is off)
        const bool band = t >= L.rtb;
        if (DevMath::any(!near && w < 1.0f && !band))
            line_image_flat_f32<DevMath, KIND>(g1, g2, g3, s, L.nv, fo, o1, o2, o3);
        if (DevMath::any(!near && w < 1.0f && band)) {
            float b1, b2, b3;
            band_image_f32<DevMath>(g1, g2, g3, t, L, b1, b2, b3);
            fo = band ? 0.0f : fo;
            o1 = band ? b1 : o1;
            o2 = band ? b2 : o2;
            o3 = band ? b3 : o3;
        }
        if (DevMath::any(!near && w > 0.0f)) {
            float c1, c2, c3;
            rule9_image_f32<DevMath>(g1, g2, g3, t, L, c1, c2, c3);
            // selects, so that a field not evaluated (or not finite
            // where its weight is zero) never enters the sum
            const bool all_c = w >= 1.0f;
            if (hard) {
                fo = all_c ? 0.0f : fo;
                o1 = all_c ? c1 : o1;
                o2 = all_c ? c2 : o2;
                o3 = all_c ? c3 : o3;
            } else {
                const float we = 1.0f - w;
                const bool mix = w > 0.0f && !all_c;
                fo = all_c ? 0.0f : (mix ? we * fo : fo);
                o1 = all_c ? c1 : (mix ? we * o1 + w * c1 : o1);
                o2 = all_c ? c2 : (mix ? we * o2 + w * c2 : o2);
                o3 = all_c ? c3 : (mix ? we * o3 + w * c3 : o3);
            }
        }
    }
    if (!near) {
        sx += (double)(L.bf[0] * fo + (o1 * L.e1f[0] + o2 * L.e2f[0] + o3 * L.nf[0]));
        sy += (double)(L.bf[1] * fo + (o1 * L.e1f[1] + o2 * L.e2f[1] + o3 * L.nf[1]));
        sz += (double)(L.bf[2] * fo + (o1 * L.e1f[2] + o2 * L.e2f[2] + o3 * L.nf[2]));
    }
}

// The images of a flat loop in the FP32 far pass, in the order of (di, dj, dk)
// with dk fastest.  A +-1 plan whose loop leaves images out (wskip: the images
// every lane of the wave leaves out) visits only the others, in the same order
// (bit 9 (di + 1) + 3 (dj + 1) + (dk + 1) rises with it).
template <int KIND, bool CUBATURE, bool SKIPS>
__device__ __forceinline__ void far_flat_loop(const FrameAtom& f, const LineLoop& L,
                                              const SegP* __restrict__ s,
                                              const SegD* __restrict__ segd, int img, int inner,
                                              unsigned lm, unsigned wskip, double& sx,
                                              double& sy, double& sz)
{
    if (SKIPS && img == 1 && L.skip_any) {
        unsigned bits = ((1u << 27) - 1u) & ~wskip;
        if (inner >= 1) bits = 0u;
        else if (inner == 0) bits &= ~(1u << 13);          // the central image
        while (bits) {
            const int b = __builtin_ctz(bits);
            bits &= bits - 1u;
            far_flat_image<KIND, CUBATURE, SKIPS>(f, L, s, segd, 1, b / 9 - 1, (b / 3) % 3 - 1,
                                                  b % 3 - 1, lm, sx, sy, sz);
        }
        return;
    }
    for (int di = -img; di <= img; di++)
        for (int dj = -img; dj <= img; dj++)
            for (int dk = -img; dk <= img; dk++) {
                if (inner_out(di, dj, dk, inner)) continue;
                far_flat_image<KIND, CUBATURE, SKIPS>(f, L, s, segd, img, di, dj, dk, lm, sx, sy,
                                                      sz);
            }
}

// PASS_FAR sums loops of one kind (KIND), each kind in its own kernel so that
// each is compiled (and its registers allocated) for its own form.  SKIPS: the
// plan leaves images out for some atoms (image skips); the kernels of plans
// that do not carry none of that work.
template <int PASS, int KIND = LOOP_3D, bool CUBATURE = false, bool SKIPS = false>
__device__ __forceinline__ bool atom_line(double px, double py, double pz,
                                          const LineLoop* __restrict__ loops, int l0, int l1,
                                          const SegF* __restrict__ segf,
                                          const SegD* __restrict__ segd, double Lx, double Ly,
                                          double Lz, int img, int inner, double& sx, double& sy,
                                          double& sz)
{
    bool any = false;
    const int oct = SKIPS && l0 < l1 ? atom_octant(px, py, pz, loops[l0]) : 0;
    unsigned present = 0u;                 // the octants of the wave's lanes
    if (SKIPS && PASS == PASS_FAR)
        for (int o = 0; o < 8; o++)
            if (DevMath::any(oct == o)) present |= 1u << o;
    for (int l = l0; l < l1; l++) {
        const LineLoop& L = loops[l];
        const double pcx = px - L.cx, pcy = py - L.cy, pcz = pz - L.cz;
        FrameAtom f;
        if (PASS == PASS_FP64) {
            if ((KIND == FP64_GLOBAL && L.flat64) || (KIND == FP64_FRAME && !L.flat64)) continue;
            if (L.flat64) frame_atom(pcx, pcy, pcz, L, f);
            const unsigned lm = SKIPS ? lane_skip(L, oct) : 0u;
            for (int di = -img; di <= img; di++)
                for (int dj = -img; dj <= img; dj++)
                    for (int dk = -img; dk <= img; dk++) {
                        if (inner_out(di, dj, dk, inner)) continue;
                        const bool skip = SKIPS ? lane_skips_image(L, lm, di, dj, dk) : false;
                        if (SKIPS && L.skip_any && DevMath::all(skip)) continue;
                        if (L.flat64) {
                            double q1, q2, q3;
                            frame_image(f, di, dj, dk, img, L, segd, q1, q2, q3);
                            add_flat_f64(q1, q2, q3, L, segd, skip, sx, sy, sz);
                            continue;
                        }
                        const double qx = pcx - di * Lx, qy = pcy - dj * Ly, qz = pcz - dk * Lz;
                        double ox = 0.0, oy = 0.0, oz = 0.0;
                        line_image_f64<DevMath>(qx, qy, qz, segd + L.seg0, L.nv, L.bx, L.by,
                                                L.bz, ox, oy, oz);
                        if (!skip) { sx += ox; sy += oy; sz += oz; }
                    }
            continue;
        }
        frame_atom(pcx, pcy, pcz, L, f);
        if (PASS == PASS_FAR && KIND != LOOP_3D) {
            const SegP* sp = reinterpret_cast<const SegP*>(segf + L.seg0);
            unsigned wskip = 0u;           // images every lane of the wave leaves out
            if (SKIPS && L.skip_any) {
                wskip = ~0u;
                for (int o = 0; o < 8; o++)
                    if ((present >> o) & 1u) wskip &= L.oct_skip[o];
            }
            far_flat_loop<KIND, CUBATURE, SKIPS>(f, L, sp, segd, img, inner,
                                                 SKIPS ? lane_skip(L, oct) : 0u, wskip, sx, sy,
                                                 sz);
            continue;
        }
        if (PASS == PASS_NEAR) {
            // only the images that can be near (usually none or one per loop)
            int i0, i1, j0, j1, k0, k1;
            near_range(pcx, L.iLx, L.reach, img, i0, i1);
            near_range(pcy, L.iLy, L.reach, img, j0, j1);
            near_range(pcz, L.iLz, L.reach, img, k0, k1);
            for (int di = i0; di <= i1; di++)
                for (int dj = j0; dj <= j1; dj++)
                    for (int dk = k0; dk <= k1; dk++) {
                        // a lane near an image never leaves it out (a left-out
                        // image lies beyond the near slab of every lane that
                        // leaves it out): only images left out for every lane
                        // are passed over
                        if (inner_out(di, dj, dk, inner) || skip_bit(L.skip_all, di, dj, dk))
                            continue;
                        double q1, q2, q3;
                        if (!frame_image(f, di, dj, dk, img, L, segd, q1, q2, q3)) continue;
                        any = true;
                        if (L.flat64) {
                            add_flat_f64(q1, q2, q3, L, segd, false, sx, sy, sz);
                            continue;
                        }
                        const double qx = pcx - di * Lx, qy = pcy - dj * Ly, qz = pcz - dk * Lz;
                        double ox = 0.0, oy = 0.0, oz = 0.0;
                        line_image_f64<DevMath>(qx, qy, qz, segd + L.seg0, L.nv, L.bx, L.by,
                                                L.bz, ox, oy, oz);
                        sx += ox; sy += oy; sz += oz;
                    }
            continue;
        }
        // PASS_FAR, LOOP_3D: the global-axes FP32 form
        const unsigned lm = SKIPS ? lane_skip(L, oct) : 0u;
        for (int di = -img; di <= img; di++)
            for (int dj = -img; dj <= img; dj++)
                for (int dk = -img; dk <= img; dk++) {
                    if (inner_out(di, dj, dk, inner)) continue;
                    const bool skip = SKIPS ? lane_skips_image(L, lm, di, dj, dk) : false;
                    if (SKIPS && L.skip_any && DevMath::all(skip)) continue;
                    double q1, q2, q3;
                    const bool near = frame_image(f, di, dj, dk, img, L, segd, q1, q2, q3);
                    const double qx = pcx - di * Lx, qy = pcy - dj * Ly, qz = pcz - dk * Lz;
                    float ox = 0.f, oy = 0.f, oz = 0.f;
                    line_image_f32<DevMath>((float)qx, (float)qy, (float)qz, segf + L.seg0, L.nv,
                                            (float)L.bx, (float)L.by, (float)L.bz, ox, oy, oz);
                    if (!near && !skip) {
                        sx += (double)ox; sy += (double)oy; sz += (double)oz;
                    }
                }
    }
    return PASS == PASS_NEAR ? any : true;
}

// Kroupa sum over loops [l0, l1); FP32: partial sums of 64 loopons in float,
// flushed to FP64.
template <int PREC>
__device__ __forceinline__ void atom_loopon(double px, double py, double pz,
                                            const LooponLoop* __restrict__ loops, int l0, int l1,
                                            const float4* __restrict__ pf,
                                            const double4* __restrict__ pd, double Lx,
                                            double Ly, double Lz, int img, int inner,
                                            double& ax, double& ay, double& az)
{
    for (int l = l0; l < l1; l++) {
        const LooponLoop& L = loops[l];
        const double pcx = px - L.cx, pcy = py - L.cy, pcz = pz - L.cz;
        double s1 = 0.0, s2 = 0.0, s3x = 0.0, s3y = 0.0, s3z = 0.0;
        if (PREC == DISCO_PREC_FP32C && L.flat) {
            // flat loop: the lane in the loop frame, loopons in the plane
            const double Q1 = pcx * L.e1x + pcy * L.e1y + pcz * L.e1z;
            const double Q2 = pcx * L.e2x + pcy * L.e2y + pcz * L.e2z;
            const double Q3 = pcx * L.ux + pcy * L.uy + pcz * L.uz;
            const float cb = (float)(L.c1 * L.b_n);
            const double r12 = L.c1 / L.c2;
            double f1 = 0.0, f2 = 0.0, f3 = 0.0;          // s3 in the frame
            for (int di = -img; di <= img; di++)
                for (int dj = -img; dj <= img; dj++)
                    for (int dk = -img; dk <= img; dk++) {
                        if (max(abs(di), max(abs(dj), abs(dk))) <= inner) continue;
                        const double q1 = Q1 - (di * L.X1 + dj * L.Y1 + dk * L.Z1);
                        const double q2 = Q2 - (di * L.X2 + dj * L.Y2 + dk * L.Z2);
                        const double q3 = Q3 - (di * L.X3 + dj * L.Y3 + dk * L.Z3);
                        const float g1 = (float)q1, g2 = (float)q2, h = (float)q3;
                        const float h2 = h * h, m3h = L.m3 * h, hn = (float)L.nmag * h;
                        const bool checked =
                            DevMath::any(q1 * q1 + q2 * q2 + q3 * q3 <= L.rmax2);
                        for (int c0 = 0; c0 < L.np; c0 += kLooponChunk) {
                            const int cnt = min(kLooponChunk, L.np - c0);
                            float t1 = 0.f, t2 = 0.f, a1 = 0.f, a2 = 0.f, ac = 0.f;
                            if (checked)
                                loopon_flat_f32<DevMath, true>(g1, g2, h2, pf + L.p0 + c0, cnt,
                                                               L.m1, L.m2, m3h, hn, cb, t1, t2,
                                                               a1, a2, ac);
                            else
                                loopon_flat_f32<DevMath, false>(g1, g2, h2, pf + L.p0 + c0, cnt,
                                                                L.m1, L.m2, m3h, hn, cb, t1, t2,
                                                                a1, a2, ac);
                            s1 += r12 * t1;
                            s2 += L.c1 * (double)(hn * t2);
                            f1 += a1; f2 += a2; f3 += (double)(h * ac);
                        }
                    }
            s3x = f1 * L.e1x + f2 * L.e2x + f3 * L.ux;
            s3y = f1 * L.e1y + f2 * L.e2y + f3 * L.uy;
            s3z = f1 * L.e1z + f2 * L.e2z + f3 * L.uz;
            ax += L.bx * (L.nx * s1 + L.mx * s2 + s3x);
            ay += L.by * (L.ny * s1 + L.my * s2 + s3y);
            az += L.bz * (L.nz * s1 + L.mz * s2 + s3z);
            continue;
        }
        for (int di = -img; di <= img; di++)
            for (int dj = -img; dj <= img; dj++)
                for (int dk = -img; dk <= img; dk++) {
                    if (max(abs(di), max(abs(dj), abs(dk))) <= inner) continue;   // inner shell
                    const double qx = pcx - di * Lx, qy = pcy - dj * Ly, qz = pcz - dk * Lz;
                    if (PREC == DISCO_PREC_FP64) {
                        loopon_image_f64<DevMath>(qx, qy, qz, pd + L.p0, L.np, L.mx, L.my, L.mz,
                                                  L.nx, L.ny, L.nz, L.b_n, L.c1, L.c2, s1, s2,
                                                  s3x, s3y, s3z);
                    } else {
                        const float fx = (float)qx, fy = (float)qy, fz = (float)qz;
                        const float c2 = (float)L.c2;
                        const float cb = (float)(L.c1 * L.b_n);
                        const double r12 = L.c1 / L.c2;
                        // wave-uniform: can any lane sit on a loopon of this image?
                        const bool checked = DevMath::any(qx * qx + qy * qy + qz * qz <= L.rmax2);
                        for (int c0 = 0; c0 < L.np; c0 += kLooponChunk) {
                            const int cnt = min(kLooponChunk, L.np - c0);
                            float t1 = 0.f, t2 = 0.f, t3x = 0.f, t3y = 0.f, t3z = 0.f;
                            if (checked)
                                loopon_image_f32<DevMath, true>(
                                    fx, fy, fz, pf + L.p0 + c0, cnt, c2 * (float)L.mx,
                                    c2 * (float)L.my, c2 * (float)L.mz, (float)L.nx, (float)L.ny,
                                    (float)L.nz, cb, t1, t2, t3x, t3y, t3z);
                            else
                                loopon_image_f32<DevMath, false>(
                                    fx, fy, fz, pf + L.p0 + c0, cnt, c2 * (float)L.mx,
                                    c2 * (float)L.my, c2 * (float)L.mz, (float)L.nx, (float)L.ny,
                                    (float)L.nz, cb, t1, t2, t3x, t3y, t3z);
                            s1 += r12 * t1; s2 += L.c1 * t2;
                            s3x += t3x; s3y += t3y; s3z += t3z;
                        }
                    }
                }
        ax += L.bx * (L.nx * s1 + L.mx * s2 + s3x);
        ay += L.by * (L.ny * s1 + L.my * s2 + s3y);
        az += L.bz * (L.nz * s1 + L.mz * s2 + s3z);
    }
}

// ------------------------------------------------------------------ one-sample kernels
// FP64 line sum (PASS_FP64) or the FP32C far pass over loops of kind KIND
// (PASS_FAR): write u (or add).  Waves per SIMD: the FP64 loop-frame walk,
// two segments per iteration, takes 5 (102 registers: the compiler then keeps
// a few values of the rare branches and of the per-loop set-up in scratch,
// none of the walk's; 1,490 ms against 1,510 at 4 waves and 1,561 at 3 on the
// 4M-atom sample, the same bits); the global-axes FP64 walk keeps 3 (168
// registers, no scratch); the FP32 ones use at most 74 registers either way.
template <int PASS, int KIND, bool CUBATURE = false, bool SKIPS = false>
__global__
__launch_bounds__(kBlock, (PASS == PASS_FP64 && KIND == FP64_FRAME) ? 5 : 3) void line_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    double* __restrict__ ux, double* __restrict__ uy, double* __restrict__ uz, int n,
    const LineLoop* __restrict__ loops, int nloops, const SegF* __restrict__ segf,
    const SegD* __restrict__ segd, double Lx, double Ly, double Lz, int img, int inner,
    int accumulate)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    atom_line<PASS, KIND, CUBATURE, SKIPS>(x[i], y[i], z[i], loops, 0, nloops, segf, segd, Lx,
                                           Ly, Lz, img, inner, sx, sy, sz);
    if (accumulate) { sx += ux[i]; sy += uy[i]; sz += uz[i]; }
    ux[i] = sx; uy[i] = sy; uz[i] = sz;
}

// FP32C near pass: adds the FP64 field of the near (loop, image) pairs.  At
// least 3 waves per SIMD (168 registers; a few values of the set-up in
// scratch): 3.0 ms against 4.0 at the 2 waves its 174 registers give, on the
// 200-loop +-4 sample's near slabs, the same bits.
__global__ __launch_bounds__(kBlock, 3) void line_near_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    double* __restrict__ ux, double* __restrict__ uy, double* __restrict__ uz, int n,
    const LineLoop* __restrict__ loops, int nloops, const SegD* __restrict__ segd, double Lx,
    double Ly, double Lz, int img, int inner)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    if (atom_line<PASS_NEAR>(x[i], y[i], z[i], loops, 0, nloops, nullptr, segd, Lx, Ly, Lz, img,
                             inner, sx, sy, sz)) {
        ux[i] += sx; uy[i] += sy; uz[i] += sz;
    }
}

// The FP32 loopon walk fits 64 registers: 8 waves per SIMD (81 registers and 6
// waves without the bound; 200.7 -> 182.4 ms on the 432,000-atom loopon sample,
// the same bits).
template <int PREC>
__global__ __launch_bounds__(kBlock, PREC == DISCO_PREC_FP32C ? 8 : 1) void loopon_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    double* __restrict__ ux, double* __restrict__ uy, double* __restrict__ uz, int n,
    const LooponLoop* __restrict__ loops, int nloops, const float4* __restrict__ pf,
    const double4* __restrict__ pd, double Lx, double Ly, double Lz, int img, int inner,
    int accumulate)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double ax = 0.0, ay = 0.0, az = 0.0;
    atom_loopon<PREC>(x[i], y[i], z[i], loops, 0, nloops, pf, pd, Lx, Ly, Lz, img, inner, ax, ay,
                      az);
    if (accumulate) { ax += ux[i]; ay += uy[i]; az += uz[i]; }
    ux[i] = ax; uy[i] = ay; uz[i] = az;
}

// A loopon plan over few atoms (the nodes of a coarse grid): the loops split
// into `split` contiguous groups along the grid's y dimension, so the launch
// has split times as many workgroups; each group's sum goes to its own slice of
// `partial`, and loopon_split_sum_kernel adds the slices in group order.
template <int PREC>
__global__ __launch_bounds__(kBlock, PREC == DISCO_PREC_FP32C ? 8 : 1) void loopon_split_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    int n, const LooponLoop* __restrict__ loops, int nloops, const float4* __restrict__ pf,
    const double4* __restrict__ pd, double Lx, double Ly, double Lz, int img, int inner,
    int split, double* __restrict__ partial)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int g = blockIdx.y;
    const int l0 = (int)((long long)g * nloops / split);
    const int l1 = (int)((long long)(g + 1) * nloops / split);
    double ax = 0.0, ay = 0.0, az = 0.0;
    atom_loopon<PREC>(x[i], y[i], z[i], loops, l0, l1, pf, pd, Lx, Ly, Lz, img, inner, ax, ay,
                      az);
    double* p = partial + (size_t)g * 3 * (size_t)n;
    p[i] = ax;
    p[(size_t)n + i] = ay;
    p[2 * (size_t)n + i] = az;
}

__global__ __launch_bounds__(kBlock) void loopon_split_sum_kernel(
    const double* __restrict__ partial, int n, int split, double* __restrict__ ux,
    double* __restrict__ uy, double* __restrict__ uz, int accumulate)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int g = 0; g < split; g++) {
        const double* p = partial + (size_t)g * 3 * (size_t)n;
        ax += p[i];
        ay += p[(size_t)n + i];
        az += p[2 * (size_t)n + i];
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

// ------------------------------------------------------------------ lattice
// Sites of a cubic lattice of nx x ny x nz cells centred on the origin, in the
// order of Code/lattice.hpp:44-48 (cells i, j, k with k fastest, then the basis
// sites), each coordinate (i a0 + f a0) - h rounded after every operation (no
// contraction), as the reference's host loop evaluates it.  Grid: (row tiles,
// ny, nx); a row is the nz * NB sites of one (i, j).
template <int NB>
__global__ __launch_bounds__(kBlock) void lattice_kernel(int ny, int nz, double a0, double hx,
                                                         double hy, double hz,
                                                         double* __restrict__ x,
                                                         double* __restrict__ y,
                                                         double* __restrict__ z)
{
#pragma clang fp contract(off)
    const int t = blockIdx.x * kBlock + threadIdx.x;       // site within the row
    if (t >= nz * NB) return;
    const int i = blockIdx.z, j = blockIdx.y, k = t / NB, b = t - k * NB;
    // basis (Code/lattice.hpp:38-39): BCC {0, (1/2,1/2,1/2)}; FCC {0, (1/2,1/2,0),
    // (1/2,0,1/2), (0,1/2,1/2)}
    const double fx = NB == 2 ? (b ? 0.5 : 0.0) : (b == 1 || b == 2 ? 0.5 : 0.0);
    const double fy = NB == 2 ? (b ? 0.5 : 0.0) : (b == 1 || b == 3 ? 0.5 : 0.0);
    const double fz = NB == 2 ? (b ? 0.5 : 0.0) : (b == 2 || b == 3 ? 0.5 : 0.0);
    const long long o = ((long long)i * ny + j) * ((long long)nz * NB) + t;
    x[o] = ((double)i * a0 + fx * a0) - hx;
    y[o] = ((double)j * a0 + fy * a0) - hy;
    z[o] = ((double)k * a0 + fz * a0) - hz;
}

// ------------------------------------------------------------------ grid field
// Lagrange weights of the ORDER-node stencil {0, .., ORDER-1} at fractional
// position f in [0, 1) of its central interval (between nodes ORDER/2 - 1
// and ORDER/2): w_j = prod_{k != j} (f + ORDER/2 - 1 - k) / (j - k).
template <int ORDER>
__device__ __forceinline__ void lagrange_weights(double f, double w[ORDER])
{
    const double t = f + (ORDER / 2 - 1);          // position in node units
    double d[ORDER];
#pragma unroll
    for (int k = 0; k < ORDER; k++) d[k] = t - k;
#pragma unroll
    for (int j = 0; j < ORDER; j++) {
        double num = 1.0, den = 1.0;
#pragma unroll
        for (int k = 0; k < ORDER; k++)
            if (k != j) { num *= d[k]; den *= (double)(j - k); }
        w[j] = num / den;
    }
}

// 4-node cubic weights in the closed form of the cubic interpolant (kept as
// it is: the far-shell results of the 4-node grid stay bit for bit)
__device__ __forceinline__ void cubic_weights(double f, double w[4])
{
    const double fm1 = f - 1.0, fm2 = f - 2.0, fp1 = f + 1.0;
    w[0] = -f * fm1 * fm2 * (1.0 / 6.0);
    w[1] = fp1 * fm1 * fm2 * 0.5;
    w[2] = -fp1 * f * fm2 * 0.5;
    w[3] = fp1 * f * fm1 * (1.0 / 6.0);
}

// First node of the ORDER-node stencil of position t (node units) on an axis of
// nn nodes, kept inside the grid; f = t - (its central interval's start).
template <int ORDER>
__device__ __forceinline__ int stencil_start(double t, int nn, double& f)
{
    double c = floor(t);
    c = fmin(fmax(c, (double)(ORDER / 2 - 1)), (double)(nn - ORDER / 2 - 1));
    f = t - c;
    return (int)c - (ORDER / 2 - 1);
}

template <int ORDER>
__global__ __launch_bounds__(kBlock) void grid_interp_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    double* __restrict__ ux, double* __restrict__ uy, double* __restrict__ uz, int n,
    const double* __restrict__ gx, const double* __restrict__ gy, const double* __restrict__ gz,
    int nx, int ny, int nz, double ox, double oy, double oz, double ih, int accumulate)
{
    const int a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= n) return;
    double fx, fy, fz, wx[ORDER], wy[ORDER], wz[ORDER];
    const int i0 = stencil_start<ORDER>((x[a] - ox) * ih, nx, fx);
    const int j0 = stencil_start<ORDER>((y[a] - oy) * ih, ny, fy);
    const int k0 = stencil_start<ORDER>((z[a] - oz) * ih, nz, fz);
    if (ORDER == 4) {
        cubic_weights(fx, wx);
        cubic_weights(fy, wy);
        cubic_weights(fz, wz);
    } else {
        lagrange_weights<ORDER>(fx, wx);
        lagrange_weights<ORDER>(fy, wy);
        lagrange_weights<ORDER>(fz, wz);
    }
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = 0; k < ORDER; k++)
        for (int j = 0; j < ORDER; j++) {
            const double wjk = wy[j] * wz[k];
            const long long row = ((long long)(k0 + k) * ny + (j0 + j)) * nx + i0;
            double rx = 0.0, ry = 0.0, rz = 0.0;
#pragma unroll
            for (int i = 0; i < ORDER; i++) {
                rx += wx[i] * gx[row + i];
                ry += wx[i] * gy[row + i];
                rz += wx[i] * gz[row + i];
            }
            sx += wjk * rx; sy += wjk * ry; sz += wjk * rz;
        }
    if (accumulate) { sx += ux[a]; sy += uy[a]; sz += uz[a]; }
    ux[a] = sx; uy[a] = sy; uz[a] = sz;
}

// ------------------------------------------------------------------ batched samples
// One sample of a batch: its atoms [a0, a0 + n) of the concatenated arrays,
// its loops [kofs[0], kofs[4]) of the concatenated loop table, those of kind k
// being [kofs[k], kofs[k+1]), and its box.
struct alignas(64) BatchSample {
    long long a0;
    int n;
    int kofs[5];
    double Lx, Ly, Lz;
};
// A work item: atoms [start, start + count) of one sample.
struct alignas(16) BatchItem {
    int sample, start, count, pad;
};

// Work-queue counters: one per pass, cleared by a kernel (no memset node in a
// captured graph).
__global__ void reset_counters_kernel(unsigned* __restrict__ c, int n)
{
    if (threadIdx.x < n) c[threadIdx.x] = 0u;
}

// Persistent batched line pass: each workgroup takes items from the queue
// (one vector atomic per item) until every item is taken; items are sorted by
// decreasing cost on the host (longest first), so the tail is short.  The
// near pass at least 3 waves per SIMD, as line_near_kernel.
template <int BLOCK, int PASS, int KIND>
__global__ __launch_bounds__(BLOCK, PASS == PASS_NEAR ? 3 : 1) void batch_line_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    double* __restrict__ ux, double* __restrict__ uy, double* __restrict__ uz,
    const BatchSample* __restrict__ samples, const BatchItem* __restrict__ items, int nitems,
    unsigned* __restrict__ counter, const LineLoop* __restrict__ loops,
    const SegF* __restrict__ segf, const SegD* __restrict__ segd, int img, int inner,
    int accumulate)
{
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= nitems) return;
        const BatchItem I = items[it];
        const BatchSample& S = samples[I.sample];
        const int l0 = PASS == PASS_FAR ? S.kofs[KIND] : S.kofs[0];
        const int l1 = PASS == PASS_FAR ? S.kofs[KIND + 1] : S.kofs[4];
        if (l0 == l1 && (accumulate || PASS == PASS_NEAR)) continue;   // nothing to add
        for (int a = threadIdx.x; a < I.count; a += BLOCK) {
            const long long i = S.a0 + I.start + a;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            const bool any = atom_line<PASS, KIND>(x[i], y[i], z[i], loops, l0, l1, segf, segd,
                                                   S.Lx, S.Ly, S.Lz, img, inner, sx, sy, sz);
            if (PASS == PASS_NEAR) {
                if (any) { ux[i] += sx; uy[i] += sy; uz[i] += sz; }
            } else {
                if (accumulate) { sx += ux[i]; sy += uy[i]; sz += uz[i]; }
                ux[i] = sx; uy[i] = sy; uz[i] = sz;
            }
        }
    }
}

}  // namespace

// Scratch of disco_hip_plan_run_noalloc: a static device array (each device
// holds its own copy for the life of the process).  Host code never passes
// its name to a kernel: the name is a host-side stand-in, and the device
// address of this device's copy comes from hipGetSymbolAddress, taken once per
// device by disco_hip_open and kept in noalloc_scratch_addr.
static __device__ __align__(256) double scratch[DISCO_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// The device address of `device`'s copy of scratch; `device` must be current.
hipError_t noalloc_scratch_addr(int device, void** out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[device] = p;
    }
    *out = g_scratch_addr[device];
    return hipSuccess;
}
}  // namespace

struct disco_plan {
    disco_ctx* ctx = nullptr;
    int kind = 0;          // 0 line, 1 loopon
    int prec = 0;
    int nloops = 0;        // packed (non-degenerate) loops
    int img = 1;
    int inner = -1;        // images with max(|i|,|j|,|k|) <= inner are skipped
    double L[3] = {0, 0, 0};
    double work = 0;       // segments (or loopons) per image
    int kofs[5] = {0, 0, 0, 0, 0};   // line plans: loops of kind k are [kofs[k], kofs[k+1])
    bool skips = false;              // line plans: some loop leaves images out (image skips)
    bool cubature = false;           // FP32C line plans: the far pass weighs in the cubature
    std::vector<LineLoop> hloops;    // line plans: host copy of the packed loops
    std::vector<int> src;            // input index of each packed loop
    int split = 1;                   // loopon plans: loop groups per launch (few atoms)
    double* d_split = nullptr;       // their partial sums (3 n split doubles, grow-only)
    size_t cap_split = 0;
    void* d_loops = nullptr;
    void* d_f = nullptr;   // SegF / float4 table
    void* d_d = nullptr;   // SegD / double4 table
    bool profile = false;
    bool profiled_run = false;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
};

struct disco_batch {
    disco_ctx* ctx = nullptr;
    int prec = 0, img = 1, nsamples = 0;
    bool kind_used[4] = {false, false, false, false};   // far-pass kernels to launch
    int inner = -1;
    std::vector<BatchSample> samples;     // host copy
    std::vector<double> seg_per_sample;   // segments of each sample
    double evals = 0;
    disco_batch_tuning tuning = {256, 256, 0};
    int nitems = 0;
    void* d_loops = nullptr;
    void* d_segf = nullptr;
    void* d_segd = nullptr;
    BatchSample* d_samples = nullptr;
    BatchItem* d_items = nullptr;
    size_t cap_items = 0;
    unsigned* d_counters = nullptr;
    hipStream_t cap_stream = nullptr;     // private stream for graph capture
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* key[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    disco_batch_tuning key_tuning = {0, 0, 0};
};

namespace {

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
    if (p->d_split) (void)hipFree(p->d_split);
    delete p;
}

// The FP32C far-pass kernel of loop kind k (SK: the plan leaves images out).
using LineKernel = void (*)(const double*, const double*, const double*, double*, double*,
                            double*, int, const LineLoop*, int, const SegF*, const SegD*, double,
                            double, double, int, int, int);
template <bool SK>
static LineKernel far_kernel(int k, bool cub)
{
    switch (k) {
    case LOOP_3D: return line_kernel<PASS_FAR, LOOP_3D, false, SK>;
    case LOOP_FLAT:
        return cub ? line_kernel<PASS_FAR, LOOP_FLAT, true, SK>
                   : line_kernel<PASS_FAR, LOOP_FLAT, false, SK>;
    case LOOP_GLIDE:
        return cub ? line_kernel<PASS_FAR, LOOP_GLIDE, true, SK>
                   : line_kernel<PASS_FAR, LOOP_GLIDE, false, SK>;
    default:
        return cub ? line_kernel<PASS_FAR, LOOP_PRISM, true, SK>
                   : line_kernel<PASS_FAR, LOOP_PRISM, false, SK>;
    }
}

// scratch = nullptr: a split loopon plan keeps its partial sums in its own
// buffer (allocated, grow-only); otherwise they go to `scratch`, which holds
// scratch_bytes (a larger need fails before anything is launched).  la, lb:
// only the packed loops [la, lb) are summed (lb < 0: all of them; a range is
// summed unsplit).
int launch(disco_plan* p, int n, const double* x, const double* y, const double* z, double* ux,
           double* uy, double* uz, int accumulate, hipStream_t s, double* scratch = nullptr,
           size_t scratch_bytes = 0, int la = 0, int lb = -1)
{
    disco_ctx* ctx = p->ctx;
    if (n == 0) return DISCO_OK;
    if (lb < 0) lb = p->nloops;
    const bool ranged = la != 0 || lb != p->nloops;
    const int nl = lb - la;
    const int split = ranged ? 1 : (p->split < p->nloops ? p->split : p->nloops);
    const size_t split_need = (size_t)3 * (size_t)n * (size_t)split * sizeof(double);
    if (p->kind == 1 && split > 1 && scratch && split_need > scratch_bytes)
        return fail(ctx, DISCO_ERR_TOO_LARGE,
                    "partial sums of " + std::to_string(split_need) + " bytes (24 B x atoms x " +
                        "loop groups) exceed the " + std::to_string(scratch_bytes) +
                        "-byte static scratch");
    const dim3 grid((unsigned)((n + kBlock - 1) / kBlock)), block(kBlock);
    const bool prof = p->profile && nl > 0 && !ranged;
    if (!ranged) p->profiled_run = false;
    if (prof) DISCO_HIP_TRY(ctx, hipEventRecord(p->ev[0], s));
    if (nl <= 0) {
        if (!accumulate) hipLaunchKernelGGL(zero3_kernel, grid, block, 0, s, ux, uy, uz, n);
        DISCO_HIP_TRY(ctx, hipGetLastError());
        return DISCO_OK;
    }
    if (p->kind == 0) {
        const LineLoop* lp = (const LineLoop*)p->d_loops;
        const SegF* sf = (const SegF*)p->d_f;
        const SegD* sd = (const SegD*)p->d_d;
        if (p->prec == DISCO_PREC_FP64) {
            // loops in global axes, then the loop-frame ones (each kernel
            // skips the others); the first writes u (or adds, as asked).
            // Both always run, so a batch sums in the same order.
            auto kg = p->skips ? line_kernel<PASS_FP64, FP64_GLOBAL, false, true>
                               : line_kernel<PASS_FP64, FP64_GLOBAL>;
            auto kf = p->skips ? line_kernel<PASS_FP64, FP64_FRAME, false, true>
                               : line_kernel<PASS_FP64, FP64_FRAME>;
            hipLaunchKernelGGL(kg, grid, block, 0, s, x, y, z, ux, uy, uz, n, lp + la, nl, sf, sd,
                               p->L[0], p->L[1], p->L[2], p->img, p->inner, accumulate);
            hipLaunchKernelGGL(kf, grid, block, 0, s, x, y, z, ux, uy, uz, n, lp + la, nl, sf, sd,
                               p->L[0], p->L[1], p->L[2], p->img, p->inner, 1);
        } else {
            // far pass: one kernel per loop kind present; the first writes u
            // (or adds, as asked), the others add
            int acc = accumulate;
            for (int k = 0; k < 4; k++) {
                const int l0 = std::max(p->kofs[k], la);
                const int cnt = std::min(p->kofs[k + 1], lb) - l0;
                if (cnt <= 0) continue;
                auto kern = p->skips ? far_kernel<true>(k, p->cubature)
                                     : far_kernel<false>(k, p->cubature);
                hipLaunchKernelGGL(kern, grid, block, 0, s, x, y, z, ux, uy, uz, n, lp + l0, cnt,
                                   sf, sd, p->L[0], p->L[1], p->L[2], p->img, p->inner, acc);
                acc = 1;
            }
            if (prof) DISCO_HIP_TRY(ctx, hipEventRecord(p->ev[1], s));
            hipLaunchKernelGGL(line_near_kernel, grid, block, 0, s, x, y, z, ux, uy, uz, n,
                               lp + la, nl, sd, p->L[0], p->L[1], p->L[2], p->img, p->inner);
        }
    } else {
        const LooponLoop* lp = (const LooponLoop*)p->d_loops;
        const float4* pf = (const float4*)p->d_f;
        const double4* pd = (const double4*)p->d_d;
        if (split > 1) {
            double* partial = scratch;
            if (!partial) {
                if (p->cap_split < split_need) {
                    if (p->d_split) DISCO_HIP_TRY(ctx, hipFree(p->d_split));
                    p->d_split = nullptr;
                    p->cap_split = 0;
                    DISCO_HIP_TRY(ctx, hipMalloc(&p->d_split, split_need));
                    p->cap_split = split_need;
                }
                partial = p->d_split;
            }
            const dim3 g2(grid.x, (unsigned)split);
            if (p->prec == DISCO_PREC_FP64)
                hipLaunchKernelGGL(loopon_split_kernel<DISCO_PREC_FP64>, g2, block, 0, s, x, y, z,
                                   n, lp, p->nloops, pf, pd, p->L[0], p->L[1], p->L[2], p->img,
                                   p->inner, split, partial);
            else
                hipLaunchKernelGGL(loopon_split_kernel<DISCO_PREC_FP32C>, g2, block, 0, s, x, y,
                                   z, n, lp, p->nloops, pf, pd, p->L[0], p->L[1], p->L[2], p->img,
                                   p->inner, split, partial);
            DISCO_HIP_TRY(ctx, hipGetLastError());
            hipLaunchKernelGGL(loopon_split_sum_kernel, grid, block, 0, s, partial, n, split,
                               ux, uy, uz, accumulate);
        } else if (p->prec == DISCO_PREC_FP64)
            hipLaunchKernelGGL(loopon_kernel<DISCO_PREC_FP64>, grid, block, 0, s, x, y, z, ux, uy,
                               uz, n, lp + la, nl, pf, pd, p->L[0], p->L[1], p->L[2], p->img,
                               p->inner, accumulate);
        else
            hipLaunchKernelGGL(loopon_kernel<DISCO_PREC_FP32C>, grid, block, 0, s, x, y, z, ux,
                               uy, uz, n, lp + la, nl, pf, pd, p->L[0], p->L[1], p->L[2],
                               p->img, p->inner, accumulate);
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
    const std::string e = pack_line(loops, nloops, Lx, Ly, Lz, pk);
    if (!e.empty()) return fail(ctx, DISCO_ERR_ARG, e);
    const double nimg = (2.0 * img + 1) * (2.0 * img + 1) * (2.0 * img + 1);
    if ((double)pk.loops.size() * nimg * 32.0 > kMaxShiftBytes)
        return fail(ctx, DISCO_ERR_TOO_LARGE,
                    "image shift table (32 B x loops x (2 img + 1)^3) above 1 GiB");
    pack_shifts(pk, img);
    disco_plan* p = new (std::nothrow) disco_plan();
    if (!p) return fail(ctx, DISCO_ERR_OOM, "host allocation");
    p->ctx = ctx; p->kind = 0; p->prec = prec; p->img = img;
    p->L[0] = Lx; p->L[1] = Ly; p->L[2] = Lz;
    p->nloops = (int)pk.loops.size();
    p->work = pk.segments;
    p->hloops = pk.loops;
    for (const LineLoop& L : pk.loops) p->src.push_back(L.src);
    std::memcpy(p->kofs, pk.kofs, sizeof(p->kofs));
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
    const std::string e = pack_loopon(loops, nloops, Lx, Ly, Lz, pk);
    if (!e.empty()) return fail(ctx, DISCO_ERR_ARG, e);
    disco_plan* p = new (std::nothrow) disco_plan();
    if (!p) return fail(ctx, DISCO_ERR_OOM, "host allocation");
    p->ctx = ctx; p->kind = 1; p->prec = prec; p->img = img;
    p->L[0] = Lx; p->L[1] = Ly; p->L[2] = Lz;
    p->nloops = (int)pk.loops.size();
    p->work = pk.loopons;
    p->src = pk.src;
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
int host_sum(disco_ctx* ctx, disco_plan* p, int n,
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
    // u is accumulated on the device: the caller's u goes up with the atoms,
    // every launch adds to it, and the result comes straight back into u
    const bool any_work = n > 0 && p->nloops > 0;
    if (any_work) {
        const double* src[6] = {x, y, z, ux, uy, uz};
        double* dst[6] = {dx, dy, dz, dux, duy, duz};
        for (int a = 0; a < 6; a++)
            DISCO_HIP_TRY(ctx, hipMemcpyAsync(dst[a], src[a], bytes, hipMemcpyHostToDevice,
                                              ctx->stream));
        DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    const double t1 = now_s();
    // with per-loop times asked for, one launch per loop timed by its own events
    // (the reference's kroupaGPU launches and times each loop:
    // Code/backend_cuda.cu:643-700); u then sums loop by loop, as kroupaCPU does
    const bool per_loop = t && t->loop_sec && any_work;
    std::vector<hipEvent_t> lev;
    if (per_loop) {
        lev.assign((size_t)p->nloops + 1, nullptr);
        for (hipEvent_t& e : lev) {
            const hipError_t er = hipEventCreate(&e);
            if (er != hipSuccess) {
                for (hipEvent_t& f : lev) if (f) (void)hipEventDestroy(f);
                return fail(ctx, DISCO_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(er));
            }
        }
    }
    auto destroy = [&]() { for (hipEvent_t& e : lev) if (e) (void)hipEventDestroy(e); };
    DISCO_HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    int rc = DISCO_OK;
    if (per_loop) {
        if (hipEventRecord(lev[0], ctx->stream) != hipSuccess) rc = fail(ctx, DISCO_ERR_HIP, "hipEventRecord");
        for (int j = 0; j < p->nloops && rc == DISCO_OK; j++) {
            rc = launch(p, n, dx, dy, dz, dux, duy, duz, 1, ctx->stream, nullptr, 0, j, j + 1);
            if (rc == DISCO_OK && hipEventRecord(lev[j + 1], ctx->stream) != hipSuccess)
                rc = fail(ctx, DISCO_ERR_HIP, "hipEventRecord");
        }
    } else if (any_work) {
        rc = launch(p, n, dx, dy, dz, dux, duy, duz, 1, ctx->stream);
    }
    if (rc != DISCO_OK) { destroy(); return rc; }
    if (hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        destroy();
        return fail(ctx, DISCO_ERR_HIP, "kernel completion");
    }
    const double t2 = now_s();
    float ms = 0.f;
    std::vector<double> lsec;
    if (per_loop) {
        lsec.assign((size_t)nloops_in, 0.0);
        for (int j = 0; j < p->nloops; j++) {
            float lm = 0.f;
            if (hipEventElapsedTime(&lm, lev[j], lev[j + 1]) == hipSuccess && p->src[j] < nloops_in)
                lsec[(size_t)p->src[j]] = lm * 1e-3;
        }
    }
    destroy();
    DISCO_HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (any_work) {
        DISCO_HIP_TRY(ctx, hipMemcpyAsync(ux, dux, bytes, hipMemcpyDeviceToHost, ctx->stream));
        DISCO_HIP_TRY(ctx, hipMemcpyAsync(uy, duy, bytes, hipMemcpyDeviceToHost, ctx->stream));
        DISCO_HIP_TRY(ctx, hipMemcpyAsync(uz, duz, bytes, hipMemcpyDeviceToHost, ctx->stream));
        DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    }
    const double t3 = now_s();
    if (t) {
        t->upload = t1 - t0;
        t->kernels = t2 - t1;
        t->download = t3 - t2;
        t->kernel_ms = ms;
        if (t->loop_sec)
            for (int l = 0; l < nloops_in; l++) t->loop_sec[l] = per_loop ? lsec[(size_t)l] : 0.0;
    }
    return DISCO_OK;
}

// ------------------------------------------------------------------ batches
void batch_release_graph(disco_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void free_batch(disco_batch* b)
{
    if (!b) return;
    batch_release_graph(b);
    for (void* q : {b->d_loops, b->d_segf, b->d_segd, (void*)b->d_samples, (void*)b->d_items,
                    (void*)b->d_counters})
        if (q) (void)hipFree(q);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    delete b;
}

bool tuning_valid(const disco_batch_tuning& t)
{
    return (t.block == 64 || t.block == 128 || t.block == 256) && t.tile >= t.block &&
           t.tile % t.block == 0 && t.groups >= 1;
}

// Work items of the given samples (every `stride`-th), tiles of `tile` atoms,
// longest first.  Returns the item list and the evaluations it covers.
std::vector<BatchItem> make_items(const disco_batch* b, int tile, int stride, double* evals)
{
    struct Keyed { double cost; BatchItem it; };
    std::vector<Keyed> v;
    const double o = 2.0 * b->img + 1.0;
    const double ii = b->inner >= 0 ? 2.0 * b->inner + 1.0 : 0.0;
    const double im = o * o * o - ii * ii * ii;
    double ev = 0;
    for (int k = 0; k < b->nsamples; k += stride) {
        const BatchSample& S = b->samples[k];
        for (int st = 0; st < S.n; st += tile) {
            BatchItem it;
            it.sample = k; it.start = st; it.count = std::min(tile, S.n - st); it.pad = 0;
            const double c = (double)it.count * b->seg_per_sample[k] * im;
            v.push_back({c, it});
            ev += c;
        }
    }
    std::stable_sort(v.begin(), v.end(), [](const Keyed& a, const Keyed& c) { return a.cost > c.cost; });
    std::vector<BatchItem> out(v.size());
    for (size_t i = 0; i < v.size(); i++) out[i] = v[i].it;
    if (evals) *evals = ev;
    return out;
}

int batch_upload_items(disco_batch* b, const std::vector<BatchItem>& items)
{
    disco_ctx* ctx = b->ctx;
    batch_release_graph(b);      // a captured graph holds d_items and nitems
    if (items.size() > b->cap_items) {
        if (b->d_items) DISCO_HIP_TRY(ctx, hipFree(b->d_items));
        b->d_items = nullptr;
        b->cap_items = 0;
        DISCO_HIP_TRY(ctx, hipMalloc(&b->d_items, items.size() * sizeof(BatchItem)));
        b->cap_items = items.size();
    }
    if (!items.empty())
        DISCO_HIP_TRY(ctx, hipMemcpy(b->d_items, items.data(), items.size() * sizeof(BatchItem),
                                     hipMemcpyHostToDevice));
    b->nitems = (int)items.size();
    return DISCO_OK;
}

template <int BLOCK>
void batch_launch_block(const disco_batch* b, int groups, hipStream_t s, const double* x,
                        const double* y, const double* z, double* ux, double* uy, double* uz,
                        int accumulate)
{
    const LineLoop* lp = (const LineLoop*)b->d_loops;
    const SegF* sf = (const SegF*)b->d_segf;
    const SegD* sd = (const SegD*)b->d_segd;
    const dim3 grid((unsigned)groups), block(BLOCK);
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, b->d_counters, 5);
    if (b->prec == DISCO_PREC_FP64) {
        // loops in global axes (queue counter 0), then the loop-frame ones
        // (counter 1), as a single-sample plan sums them
        hipLaunchKernelGGL((batch_line_kernel<BLOCK, PASS_FP64, FP64_GLOBAL>), grid, block, 0, s,
                           x, y, z, ux, uy, uz, b->d_samples, b->d_items, b->nitems,
                           b->d_counters, lp, sf, sd, b->img, b->inner, accumulate);
        hipLaunchKernelGGL((batch_line_kernel<BLOCK, PASS_FP64, FP64_FRAME>), grid, block, 0, s,
                           x, y, z, ux, uy, uz, b->d_samples, b->d_items, b->nitems,
                           b->d_counters + 1, lp, sf, sd, b->img, b->inner, 1);
    } else {
        // far pass: one launch per loop kind present (queue counter k); the
        // first writes u (or adds, as asked) for every sample, the others add
        int acc = accumulate;
        bool first = true;
        for (int k = 0; k < 4; k++) {
            if (!b->kind_used[k] && !(first && k == 3)) continue;
            auto kern = k == LOOP_3D ? batch_line_kernel<BLOCK, PASS_FAR, LOOP_3D>
                      : k == LOOP_FLAT ? batch_line_kernel<BLOCK, PASS_FAR, LOOP_FLAT>
                      : k == LOOP_GLIDE ? batch_line_kernel<BLOCK, PASS_FAR, LOOP_GLIDE>
                      : batch_line_kernel<BLOCK, PASS_FAR, LOOP_PRISM>;
            hipLaunchKernelGGL(kern, grid, block, 0, s, x, y, z, ux, uy, uz, b->d_samples,
                               b->d_items, b->nitems, b->d_counters + k, lp, sf, sd, b->img,
                               b->inner, acc);
            acc = 1;
            first = false;
        }
        hipLaunchKernelGGL((batch_line_kernel<BLOCK, PASS_NEAR, LOOP_3D>), grid, block, 0, s, x,
                           y, z, ux, uy, uz, b->d_samples, b->d_items, b->nitems,
                           b->d_counters + 4, lp, sf, sd, b->img, b->inner, accumulate);
    }
}

int batch_launch(disco_batch* b, const disco_batch_tuning& t, hipStream_t s, const double* x,
                 const double* y, const double* z, double* ux, double* uy, double* uz,
                 int accumulate)
{
    if (b->nitems == 0) {
        // nothing to sum: u = 0 (or unchanged) over every sample's atoms
        if (!accumulate)
            for (const BatchSample& S : b->samples)
                if (S.n > 0)
                    hipLaunchKernelGGL(zero3_kernel, dim3((S.n + kBlock - 1) / kBlock),
                                       dim3(kBlock), 0, s, ux + S.a0, uy + S.a0, uz + S.a0, S.n);
    } else if (t.block == 64) {
        batch_launch_block<64>(b, t.groups, s, x, y, z, ux, uy, uz, accumulate);
    } else if (t.block == 128) {
        batch_launch_block<128>(b, t.groups, s, x, y, z, ux, uy, uz, accumulate);
    } else {
        batch_launch_block<256>(b, t.groups, s, x, y, z, ux, uy, uz, accumulate);
    }
    DISCO_HIP_TRY(b->ctx, hipGetLastError());
    return DISCO_OK;
}

int default_groups(disco_ctx* ctx)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) return 1024;
    return prop.multiProcessorCount * 4;
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
        case DISCO_ERR_IO: return "file could not be written";
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

int disco_hip_grid_interp(disco_ctx* ctx, int n, const double* d_x, const double* d_y,
                          const double* d_z, double* d_ux, double* d_uy, double* d_uz,
                          const double* d_gx, const double* d_gy, const double* d_gz, int nx,
                          int ny, int nz, const double* origin, double h, int order,
                          int accumulate, void* stream)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (n < 0) return fail(ctx, DISCO_ERR_ARG, "negative size");
    if (order != 4 && order != 6) return fail(ctx, DISCO_ERR_ARG, "stencil order must be 4 or 6");
    if (nx < order || ny < order || nz < order)
        return fail(ctx, DISCO_ERR_ARG, "a grid needs at least `order` nodes per axis");
    if (!origin || !(h > 0.0) || !std::isfinite(h))
        return fail(ctx, DISCO_ERR_ARG, "bad grid origin or spacing");
    if (n == 0) return DISCO_OK;
    if (!d_x || !d_y || !d_z || !d_ux || !d_uy || !d_uz || !d_gx || !d_gy || !d_gz)
        return fail(ctx, DISCO_ERR_ARG, "null device array");
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    auto kern = order == 6 ? grid_interp_kernel<6> : grid_interp_kernel<4>;
    hipLaunchKernelGGL(kern, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, d_x, d_y, d_z, d_ux, d_uy, d_uz, n, d_gx, d_gy, d_gz,
                       nx, ny, nz, origin[0], origin[1], origin[2], 1.0 / h, accumulate);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    return DISCO_OK;
}

long long disco_hip_lattice_atoms(int lattice, int nx, int ny, int nz)
{
    if ((lattice != DISCO_LATTICE_BCC && lattice != DISCO_LATTICE_FCC) || nx < 0 || ny < 0 ||
        nz < 0)
        return -1;
    return (long long)(lattice == DISCO_LATTICE_FCC ? 4 : 2) * nx * ny * nz;
}

int disco_hip_lattice(disco_ctx* ctx, int lattice, int nx, int ny, int nz, double a0,
                      double* d_x, double* d_y, double* d_z, void* stream)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (lattice != DISCO_LATTICE_BCC && lattice != DISCO_LATTICE_FCC)
        return fail(ctx, DISCO_ERR_ARG, "unknown lattice");
    if (nx < 0 || ny < 0 || nz < 0) return fail(ctx, DISCO_ERR_ARG, "negative cell count");
    if (nx > 65535 || ny > 65535 || nz > 65535)
        return fail(ctx, DISCO_ERR_TOO_LARGE, "more than 65535 cells along an axis");
    if (!std::isfinite(a0)) return fail(ctx, DISCO_ERR_ARG, "non-finite lattice constant");
    if (nx == 0 || ny == 0 || nz == 0) return DISCO_OK;
    if (!d_x || !d_y || !d_z) return fail(ctx, DISCO_ERR_ARG, "null device array");
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int nb = lattice == DISCO_LATTICE_FCC ? 4 : 2;
    const double hx = nx * a0 / 2., hy = ny * a0 / 2., hz = nz * a0 / 2.;   // Code/lattice.hpp:43
    const dim3 grid((unsigned)((nz * nb + kBlock - 1) / kBlock), (unsigned)ny, (unsigned)nx);
    hipStream_t s = (hipStream_t)stream;
    if (nb == 4)
        hipLaunchKernelGGL(lattice_kernel<4>, grid, dim3(kBlock), 0, s, ny, nz, a0, hx, hy, hz,
                           d_x, d_y, d_z);
    else
        hipLaunchKernelGGL(lattice_kernel<2>, grid, dim3(kBlock), 0, s, ny, nz, a0, hx, hy, hz,
                           d_x, d_y, d_z);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    return DISCO_OK;
}

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
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        noalloc_scratch_addr(device, &ctx->noalloc_scratch) != hipSuccess ||
        !ctx->noalloc_scratch) {
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
    disco::release_output_buffers(ctx);
    disco::release_prismatic_buffers(ctx);
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
    rc = make_line_plan(ctx, loops, nloops, Lx, Ly, Lz, img, precision, &p, nullptr);
    if (rc != DISCO_OK) return rc;
    rc = host_sum(ctx, p, n, x, y, z, ux, uy, uz, t, nloops);
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
    rc = make_loopon_plan(ctx, loops, nloops, Lx, Ly, Lz, img, precision, &p, nullptr);
    if (rc != DISCO_OK) return rc;
    rc = host_sum(ctx, p, n, x, y, z, ux, uy, uz, t, nloops);
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

int disco_hip_plan_run_noalloc(disco_plan* plan, int n, const double* d_x, const double* d_y,
                               const double* d_z, double* d_ux, double* d_uy, double* d_uz,
                               int accumulate, void* stream)
{
    if (!plan) return DISCO_ERR_ARG;
    disco_ctx* ctx = plan->ctx;
    if (n < 0) return fail(ctx, DISCO_ERR_ARG, "negative atom count");
    if (n > 0 && (!d_x || !d_y || !d_z || !d_ux || !d_uy || !d_uz))
        return fail(ctx, DISCO_ERR_ARG, "null device array");
    if (!ctx->noalloc_scratch) return fail(ctx, DISCO_ERR_STATE, "no static scratch address");
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    return launch(plan, n, d_x, d_y, d_z, d_ux, d_uy, d_uz, accumulate ? 1 : 0,
                  (hipStream_t)stream, (double*)ctx->noalloc_scratch,
                  (size_t)DISCO_NOALLOC_SCRATCH_BYTES);
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

namespace {
double shell_images(int img, int inner)
{
    const double o = 2.0 * img + 1.0;
    const double i = inner >= 0 ? 2.0 * std::min(inner, img) + 1.0 : 0.0;
    return o * o * o - i * i * i;
}
}  // namespace

double disco_hip_plan_evals_per_atom(const disco_plan* plan)
{
    if (!plan) return 0.0;
    return plan->work * shell_images(plan->img, plan->inner);
}

int disco_hip_plan_set_inner(disco_plan* plan, int inner)
{
    if (!plan) return DISCO_ERR_ARG;
    if (inner < -1 || inner > plan->img) return fail(plan->ctx, DISCO_ERR_ARG, "inner shell out of range");
    plan->inner = inner;
    return DISCO_OK;
}

int disco_hip_plan_set_cubature(disco_plan* plan, double k_exact, double k_rule, int* eligible)
{
    if (!plan) return DISCO_ERR_ARG;
    disco_ctx* ctx = plan->ctx;
    if (plan->kind != 0 || plan->prec != DISCO_PREC_FP32C)
        return fail(ctx, DISCO_ERR_ARG, "the cubature form is for FP32C line plans");
    if (k_exact <= 0.0) {
        plan->cubature = false;
        if (eligible) *eligible = 0;
        return DISCO_OK;
    }
    if (!(k_exact >= 2.0) || !(k_rule >= k_exact) || !std::isfinite(k_rule))
        return fail(ctx, DISCO_ERR_ARG, "cubature radii: need 2 <= k_exact <= k_rule < inf");
    int ok = 0;
    for (LineLoop& L : plan->hloops) {
        if (!L.rule_ok) continue;
        ok++;
        const double R = L.rR;
        // no lane of the near slab (|h| < hn, in-plane radius^2 < rn2) is weighted
        const double t0 = std::fmax(k_exact * k_exact * R * R, 1.0001 * (L.rn2 + L.hn * L.hn));
        const double t1 = std::fmax(k_rule * k_rule * R * R, t0);
        L.rt0 = (float)t0;
        L.rwinv = t1 > t0 * (1.0 + 1e-6) ? (float)(1.0 / (t1 - t0)) : 1e30f;
    }
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!plan->hloops.empty()) {
        DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        DISCO_HIP_TRY(ctx, hipMemcpy(plan->d_loops, plan->hloops.data(),
                                     plan->hloops.size() * sizeof(LineLoop),
                                     hipMemcpyHostToDevice));
    }
    plan->cubature = true;
    if (eligible) *eligible = ok;
    return DISCO_OK;
}

int disco_hip_plan_set_band(disco_plan* plan, double k_band, int* eligible)
{
    if (!plan) return DISCO_ERR_ARG;
    disco_ctx* ctx = plan->ctx;
    if (plan->kind != 0 || plan->prec != DISCO_PREC_FP32C)
        return fail(ctx, DISCO_ERR_ARG, "the band form is for FP32C line plans");
    if (!(k_band <= 0.0) && !(k_band >= 2.5 && std::isfinite(k_band)))   // NaN included
        return fail(ctx, DISCO_ERR_ARG, "band radius: need k_band >= 2.5");
    int ok = 0;
    for (LineLoop& L : plan->hloops) {
        L.rtb = 3.0e38f;
        if (k_band <= 0.0 || !L.band_ok) continue;
        ok++;
        const double R = L.rR;
        // no lane of the near slab takes the band rule
        L.rtb = (float)std::fmax(k_band * k_band * R * R, 1.0001 * (L.rn2 + L.hn * L.hn));
    }
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!plan->hloops.empty()) {
        DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        DISCO_HIP_TRY(ctx, hipMemcpy(plan->d_loops, plan->hloops.data(),
                                     plan->hloops.size() * sizeof(LineLoop),
                                     hipMemcpyHostToDevice));
    }
    if (eligible) *eligible = k_band > 0.0 ? ok : 0;
    return DISCO_OK;
}

// The image skips of a line plan: `per` masks per input loop (1: one for every
// atom, 8: one per octant of the planes split[3]), or none (skip = NULL).
static int set_image_skips(disco_plan* plan, const unsigned* skip, int nloops, int per,
                           const double* split)
{
    if (!plan) return DISCO_ERR_ARG;
    disco_ctx* ctx = plan->ctx;
    if (plan->kind != 0) return fail(ctx, DISCO_ERR_ARG, "image skips are for line plans");
    if (skip && nloops < 0) return fail(ctx, DISCO_ERR_ARG, "negative loop count");
    if (skip)
        for (const LineLoop& L : plan->hloops)
            if (L.src >= nloops) return fail(ctx, DISCO_ERR_ARG, "skip array shorter than the loops");
    if (split)
        for (int k = 0; k < 3; k++)
            if (std::isnan(split[k])) return fail(ctx, DISCO_ERR_ARG, "octant plane is NaN");
    for (LineLoop& L : plan->hloops) {
        unsigned all = ~0u, any = 0u;
        for (int o = 0; o < 8; o++) {
            const unsigned m = skip ? skip[(size_t)L.src * per + (per == 8 ? o : 0)] & ((1u << 27) - 1u)
                                    : 0u;
            L.oct_skip[o] = m;
            all &= m;
            any |= m;
        }
        L.skip_all = all;
        L.skip_any = any;
        L.osx = split ? split[0] : INFINITY;
        L.osy = split ? split[1] : INFINITY;
        L.osz = split ? split[2] : INFINITY;
    }
    plan->skips = false;
    for (const LineLoop& L : plan->hloops) plan->skips = plan->skips || L.skip_any != 0u;
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!plan->hloops.empty()) {
        DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        DISCO_HIP_TRY(ctx, hipMemcpy(plan->d_loops, plan->hloops.data(),
                                     plan->hloops.size() * sizeof(LineLoop),
                                     hipMemcpyHostToDevice));
    }
    return DISCO_OK;
}

int disco_hip_plan_set_image_skip(disco_plan* plan, const unsigned* skip, int nloops)
{
    return set_image_skips(plan, skip, nloops, 1, nullptr);
}

int disco_hip_plan_set_image_skip_octants(disco_plan* plan, const unsigned* skip, int nloops,
                                          const double* split)
{
    if (plan && skip && !split) return fail(plan->ctx, DISCO_ERR_ARG, "octant planes missing");
    return set_image_skips(plan, skip, nloops, 8, skip ? split : nullptr);
}

int disco_hip_line_cubature(const disco_line_loop* loops, int nloops, double* out)
{
    if (nloops < 0 || (nloops > 0 && (!loops || !out))) return DISCO_ERR_ARG;
    for (int l = 0; l < nloops; l++) {
        LinePack pk;
        const std::string e = pack_line(loops + l, 1, 1.0, 1.0, 1.0, pk);
        if (!e.empty()) return DISCO_ERR_ARG;
        double* o = out + 5 * (size_t)l;
        o[0] = o[1] = o[2] = o[3] = o[4] = 0.0;
        if (pk.loops.empty()) continue;
        const LineLoop& L = pk.loops[0];
        o[0] = L.rule_ok;
        o[1] = L.rw0;
        o[2] = L.rw1;
        o[3] = L.rrho;
        o[4] = L.rR;
    }
    return DISCO_OK;
}

int disco_hip_line_band(const disco_line_loop* loops, int nloops, double* out)
{
    if (nloops < 0 || (nloops > 0 && (!loops || !out))) return DISCO_ERR_ARG;
    for (int l = 0; l < nloops; l++) {
        LinePack pk;
        const std::string e = pack_line(loops + l, 1, 1.0, 1.0, 1.0, pk);
        if (!e.empty()) return DISCO_ERR_ARG;
        double* o = out + 6 * (size_t)l;
        for (int i = 0; i < 6; i++) o[i] = 0.0;
        if (pk.loops.empty()) continue;
        const LineLoop& L = pk.loops[0];
        o[0] = L.band_ok;
        o[1] = L.bw0;
        o[2] = L.bw1;
        o[3] = L.bw2;
        o[4] = L.brho1;
        o[5] = L.brho2;
    }
    return DISCO_OK;
}

int disco_hip_plan_set_split(disco_plan* plan, int groups)
{
    if (!plan) return DISCO_ERR_ARG;
    if (plan->kind != 1) return fail(plan->ctx, DISCO_ERR_ARG, "loop groups are for loopon plans");
    if (groups < 1 || groups > 256) return fail(plan->ctx, DISCO_ERR_ARG, "groups out of [1, 256]");
    plan->split = groups;
    return DISCO_OK;
}

int disco_hip_batch_set_inner(disco_batch* b, int inner)
{
    if (!b) return DISCO_ERR_ARG;
    if (inner < -1 || inner > b->img) return fail(b->ctx, DISCO_ERR_ARG, "inner shell out of range");
    if (inner != b->inner) {
        b->inner = inner;
        const double im = shell_images(b->img, inner);
        b->evals = 0;
        for (int k = 0; k < b->nsamples; k++) b->evals += b->samples[k].n * b->seg_per_sample[k] * im;
        return batch_upload_items(b, make_items(b, b->tuning.tile, 1, nullptr));  // costs changed
    }
    return DISCO_OK;
}

int disco_hip_plan_free(disco_plan* plan)
{
    free_plan(plan);
    return DISCO_OK;
}

int disco_hip_batch_plan(disco_ctx* ctx, const disco_sample* samples, int nsamples, int img,
                         int precision, disco_batch** out)
{
    if (!out) return fail(ctx, DISCO_ERR_ARG, "null batch pointer");
    *out = nullptr;
    int rc = check_common(ctx, 0, 1.0, 1.0, 1.0, img, precision, 0);
    if (rc != DISCO_OK) return rc;
    if (nsamples < 0 || (nsamples > 0 && !samples)) return fail(ctx, DISCO_ERR_ARG, "bad sample array");
    disco_batch* b = new (std::nothrow) disco_batch();
    if (!b) return fail(ctx, DISCO_ERR_OOM, "host allocation");
    b->ctx = ctx; b->prec = precision; b->img = img; b->nsamples = nsamples;
    LinePack all;
    const double im = (2.0 * img + 1) * (2.0 * img + 1) * (2.0 * img + 1);
    double shift_bytes = 0.0;
    for (int k = 0; k < nsamples; k++) {
        const disco_sample& S = samples[k];
        std::string why;
        if (S.n < 0 || S.atom_offset < 0) why = "negative atom range";
        else if (S.nloops < 0 || (S.nloops > 0 && !S.loops)) why = "bad loop array";
        else if (S.nloops > DISCO_MAX_LOOPS) { free_batch(b); return fail(ctx, DISCO_ERR_TOO_LARGE, "too many loops"); }
        else if (!(std::isfinite(S.L[0]) && std::isfinite(S.L[1]) && std::isfinite(S.L[2]))) why = "non-finite box";
        LinePack pk;
        if (why.empty()) why = pack_line(S.loops, S.nloops, S.L[0], S.L[1], S.L[2], pk);
        if (!why.empty()) {
            free_batch(b);
            return fail(ctx, DISCO_ERR_ARG, "sample " + std::to_string(k) + ": " + why);
        }
        shift_bytes += (double)pk.loops.size() * im * 32.0;
        if (shift_bytes > kMaxShiftBytes) {
            free_batch(b);
            return fail(ctx, DISCO_ERR_TOO_LARGE,
                        "image shift table (32 B x loops x (2 img + 1)^3) above 1 GiB");
        }
        pack_shifts(pk, img);
        BatchSample bs;
        std::memset(&bs, 0, sizeof(bs));
        bs.a0 = S.atom_offset; bs.n = S.n;
        const int loop_base = (int)all.loops.size();
        const int seg_base = (int)all.segd.size();
        for (LineLoop L : pk.loops) {
            L.seg0 += seg_base;
            L.sh0 += seg_base;
            all.loops.push_back(L);
        }
        all.segd.insert(all.segd.end(), pk.segd.begin(), pk.segd.end());
        all.segf.insert(all.segf.end(), pk.segf.begin(), pk.segf.end());
        for (int q = 0; q < 5; q++) bs.kofs[q] = loop_base + pk.kofs[q];
        for (int q = 0; q < 4; q++) b->kind_used[q] = b->kind_used[q] || pk.kofs[q + 1] > pk.kofs[q];
        bs.Lx = S.L[0]; bs.Ly = S.L[1]; bs.Lz = S.L[2];
        b->samples.push_back(bs);
        b->seg_per_sample.push_back(pk.segments);
        b->evals += (double)S.n * pk.segments * im;
    }
    disco_ctx* c = ctx;
    auto up = [&](void** dst, const void* src, size_t bytes) -> int {
        if (!bytes) return DISCO_OK;
        DISCO_HIP_TRY(c, hipMalloc(dst, bytes));
        DISCO_HIP_TRY(c, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return DISCO_OK;
    };
    if ((rc = hipSetDevice(ctx->device) == hipSuccess ? DISCO_OK : DISCO_ERR_HIP) != DISCO_OK ||
        (rc = up(&b->d_loops, all.loops.data(), all.loops.size() * sizeof(LineLoop))) != DISCO_OK ||
        (rc = up(&b->d_segf, all.segf.data(), all.segf.size() * sizeof(SegF))) != DISCO_OK ||
        (rc = up(&b->d_segd, all.segd.data(), all.segd.size() * sizeof(SegD))) != DISCO_OK ||
        (rc = up((void**)&b->d_samples, b->samples.data(), b->samples.size() * sizeof(BatchSample))) != DISCO_OK) {
        free_batch(b);
        return rc;
    }
    if (hipMalloc(&b->d_counters, 64 * sizeof(unsigned)) != hipSuccess ||
        hipStreamCreate(&b->cap_stream) != hipSuccess) {
        free_batch(b);
        return fail(ctx, DISCO_ERR_HIP, "batch buffers");
    }
    b->tuning = {256, 256, default_groups(ctx)};
    rc = batch_upload_items(b, make_items(b, b->tuning.tile, 1, nullptr));
    if (rc != DISCO_OK) { free_batch(b); return rc; }
    *out = b;
    return DISCO_OK;
}

int disco_hip_batch_set_tuning(disco_batch* b, const disco_batch_tuning* t)
{
    if (!b || !t) return DISCO_ERR_ARG;
    if (!tuning_valid(*t)) return fail(b->ctx, DISCO_ERR_ARG, "invalid launch shape");
    if (t->tile != b->tuning.tile) {
        const int rc = batch_upload_items(b, make_items(b, t->tile, 1, nullptr));
        if (rc != DISCO_OK) return rc;
    }
    b->tuning = *t;
    return DISCO_OK;
}

int disco_hip_batch_get_tuning(const disco_batch* b, disco_batch_tuning* t)
{
    if (!b || !t) return DISCO_ERR_ARG;
    *t = b->tuning;
    return DISCO_OK;
}

int disco_hip_batch_run(disco_batch* b, const double* x, const double* y, const double* z,
                        double* ux, double* uy, double* uz, int accumulate, void* stream)
{
    if (!b) return DISCO_ERR_ARG;
    disco_ctx* ctx = b->ctx;
    if (b->nsamples > 0 && (!x || !y || !z || !ux || !uy || !uz))
        return fail(ctx, DISCO_ERR_ARG, "null device array");
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const void* key[7] = {x, y, z, ux, uy, uz, (const void*)(intptr_t)(accumulate ? 1 : 0)};
    const bool same = b->exec && std::memcmp(key, b->key, sizeof(key)) == 0 &&
                      std::memcmp(&b->tuning, &b->key_tuning, sizeof(b->tuning)) == 0;
    if (!same) {
        batch_release_graph(b);
        DISCO_HIP_TRY(ctx, hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        const int rc = batch_launch(b, b->tuning, b->cap_stream, x, y, z, ux, uy, uz, accumulate);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
        if (rc != DISCO_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return fail(ctx, DISCO_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        b->graph = g;
        DISCO_HIP_TRY(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        std::memcpy(b->key, key, sizeof(key));
        b->key_tuning = b->tuning;
    }
    DISCO_HIP_TRY(ctx, hipGraphLaunch(b->exec, (hipStream_t)stream));
    return DISCO_OK;
}

int disco_hip_batch_autotune(disco_batch* b, const double* x, const double* y, const double* z,
                             double* ux, double* uy, double* uz, int stride, void* stream,
                             disco_batch_tuning* best, double* best_ms)
{
    if (!b || stride < 1) return DISCO_ERR_ARG;
    disco_ctx* ctx = b->ctx;
    if (b->nsamples > 0 && (!x || !y || !z || !ux || !uy || !uz))
        return fail(ctx, DISCO_ERR_ARG, "null device array");
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    const int cus = default_groups(ctx) / 4;
    std::vector<disco_batch_tuning> cand;
    for (int block : {128, 256})
        for (int mult : {1, 4})
            for (int g : {4, 8, 16})
                cand.push_back({block, block * mult, cus * g});
    const disco_batch_tuning saved = b->tuning;
    disco_batch_tuning win = saved;
    double win_ms = 1e300, full_evals = b->evals;
    for (const disco_batch_tuning& t : cand) {
        double ev = 0;
        int rc = batch_upload_items(b, make_items(b, t.tile, stride, &ev));
        if (rc != DISCO_OK) return rc;
        if (b->nitems == 0) continue;
        rc = batch_launch(b, t, s, x, y, z, ux, uy, uz, 0);        // warm
        if (rc != DISCO_OK) return rc;
        DISCO_HIP_TRY(ctx, hipEventRecord(ctx->ev0, s));
        rc = batch_launch(b, t, s, x, y, z, ux, uy, uz, 0);
        if (rc != DISCO_OK) return rc;
        DISCO_HIP_TRY(ctx, hipEventRecord(ctx->ev1, s));
        DISCO_HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
        float ms = 0.f;
        DISCO_HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        const double scaled = ev > 0 ? ms * full_evals / ev : ms;
        if (scaled < win_ms) { win_ms = scaled; win = t; }
    }
    const int rc = batch_upload_items(b, make_items(b, win.tile, 1, nullptr));
    if (rc != DISCO_OK) return rc;
    b->tuning = win;
    if (best) *best = win;
    if (best_ms) *best_ms = win_ms < 1e300 ? win_ms : 0.0;
    return DISCO_OK;
}

double disco_hip_batch_evals(const disco_batch* b) { return b ? b->evals : 0.0; }

int disco_hip_batch_free(disco_batch* b)
{
    if (b && b->ctx) (void)hipSetDevice(b->ctx->device);
    free_batch(b);
    return DISCO_OK;
}

}  // extern "C"
