// plan_pack.h -- host packing of loop geometry into the device tables.
//
// Line loops (the reference's LoopData with line = true, Code/backend.hpp:20-26,
// assembled at Code/main.cpp:689-711): vertices become offsets from the apex c
// in FP64; segment k = V[k-1] -> V[k] with V[nv-1] -> V[0] closing the contour
// (Code/barnett.hpp:62-64).  Zero-length segments are dropped: the reference
// skips their line terms (Code/barnett.hpp:77) and their fan triangle has
// num = 0, den >= 0, so atan2 gives 0 -- they contribute nothing.
//
// Loopon loops (line = false): loopons become offsets from their centroid, as
// the reference's --fp32 upload does (Code/backend_cuda.cu:676-685).
#pragma once

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "disco_hip.h"
#include "pair_terms.h"

namespace disco {

struct LinePack {
    std::vector<LineLoop> loops;
    std::vector<SegF> segf;
    std::vector<SegD> segd;
    double segments = 0;      // total segment count (work per atom-image)
    std::vector<double> work; // per input loop: its segment count
};

struct LooponPack {
    std::vector<LooponLoop> loops;
    std::vector<float4> pf;
    std::vector<double4> pd;
    double loopons = 0;
    std::vector<double> work;
};

// Height of the near slab: a lane whose loop image lies within this distance
// of the loop's plane (and inside its radius + this margin) is evaluated in
// FP64 in the FP32C mode.
inline double near_height(double bmag) { return std::fmax(3.0 * bmag, 6.0); }

// Returns "" on success, else the reason the input is rejected.
inline std::string pack_line(const disco_line_loop* in, int nloops, LinePack& out)
{
    out = LinePack();
    out.work.assign(nloops, 0.0);
    for (int l = 0; l < nloops; l++) {
        const disco_line_loop& lp = in[l];
        if (lp.nv < 0) return "negative vertex count";
        if (lp.nv > 0 && (!lp.vx || !lp.vy || !lp.vz)) return "null vertex array";
        // distinct consecutive vertices, relative to the apex
        std::vector<double> ax, ay, az;
        for (int k = 0; k < lp.nv; k++) {
            const double x = lp.vx[k] - lp.c[0], y = lp.vy[k] - lp.c[1], z = lp.vz[k] - lp.c[2];
            if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z))
                return "non-finite vertex";
            if (!ax.empty() && x == ax.back() && y == ay.back() && z == az.back()) continue;
            ax.push_back(x); ay.push_back(y); az.push_back(z);
        }
        while (ax.size() > 1 && ax.back() == ax[0] && ay.back() == ay[0] && az.back() == az[0]) {
            ax.pop_back(); ay.pop_back(); az.pop_back();
        }
        const int nv = (int)ax.size();
        out.work[l] = (double)lp.nv;
        if (nv < 3) continue;   // degenerate contour: its field is identically zero
        // vector area -> plane normal
        double Nx = 0, Ny = 0, Nz = 0;
        for (int k = 0; k < nv; k++) {
            const int j = (k + nv - 1) % nv;
            Nx += ay[j] * az[k] - az[j] * ay[k];
            Ny += az[j] * ax[k] - ax[j] * az[k];
            Nz += ax[j] * ay[k] - ay[j] * ax[k];
        }
        const double Nn = std::sqrt(Nx * Nx + Ny * Ny + Nz * Nz);
        if (!(Nn > 0.0)) return "polyline with zero vector area";
        const double nx = Nx / Nn, ny = Ny / Nn, nz = Nz / Nn;
        double hmax = 0.0, rmax = 0.0;
        for (int k = 0; k < nv; k++) {
            const double h = ax[k] * nx + ay[k] * ny + az[k] * nz;
            const double r2 = ax[k] * ax[k] + ay[k] * ay[k] + az[k] * az[k] - h * h;
            hmax = std::fmax(hmax, std::fabs(h));
            rmax = std::fmax(rmax, std::sqrt(std::fmax(r2, 0.0)));
        }
        if (hmax > 1e-6 * std::fmax(rmax, 1.0))
            return "polyline is not planar (the solid angle uses a flat cut)";

        LineLoop L;
        std::memset(&L, 0, sizeof(L));
        L.cx = lp.c[0]; L.cy = lp.c[1]; L.cz = lp.c[2];
        L.bx = lp.b[0]; L.by = lp.b[1]; L.bz = lp.b[2];
        L.nx = nx; L.ny = ny; L.nz = nz;
        const double bmag = std::sqrt(lp.b[0] * lp.b[0] + lp.b[1] * lp.b[1] + lp.b[2] * lp.b[2]);
        const double H = near_height(bmag);
        L.hn = H + hmax;
        L.rn2 = (rmax + H) * (rmax + H);
        L.seg0 = (int)out.segd.size();
        L.nv = nv;
        for (int k = 0; k < nv; k++) {
            const int j = (k + nv - 1) % nv;
            SegD d;
            std::memset(&d, 0, sizeof(d));
            d.bx = ax[k]; d.by = ay[k]; d.bz = az[k];
            const double ex = ax[k] - ax[j], ey = ay[k] - ay[j], ez = az[k] - az[j];
            const double len = std::sqrt(ex * ex + ey * ey + ez * ez);
            d.L = len;
            d.tx = ex / len; d.ty = ey / len; d.tz = ez / len;
            d.nx = ay[j] * az[k] - az[j] * ay[k];
            d.ny = az[j] * ax[k] - ax[j] * az[k];
            d.nz = ax[j] * ay[k] - ay[j] * ax[k];
            const double wx = lp.b[1] * d.tz - lp.b[2] * d.ty;
            const double wy = lp.b[2] * d.tx - lp.b[0] * d.tz;
            const double wz = lp.b[0] * d.ty - lp.b[1] * d.tx;
            d.wkx = kC3 * wx; d.wky = kC3 * wy; d.wkz = kC3 * wz;
            d.wjx = kCJ * wx; d.wjy = kCJ * wy; d.wjz = kCJ * wz;
            SegF f;
            std::memset(&f, 0, sizeof(f));
            f.bx = (float)d.bx; f.by = (float)d.by; f.bz = (float)d.bz;
            f.tx = (float)d.tx; f.ty = (float)d.ty; f.tz = (float)d.tz;
            f.L = (float)d.L;
            f.nx = (float)d.nx; f.ny = (float)d.ny; f.nz = (float)d.nz;
            f.wkx = (float)d.wkx; f.wky = (float)d.wky; f.wkz = (float)d.wkz;
            f.wjx = (float)d.wjx; f.wjy = (float)d.wjy; f.wjz = (float)d.wjz;
            out.segd.push_back(d);
            out.segf.push_back(f);
        }
        out.segments += nv;
        out.loops.push_back(L);
    }
    return "";
}

inline std::string pack_loopon(const disco_loopon_loop* in, int nloops, LooponPack& out)
{
    out = LooponPack();
    out.work.assign(nloops, 0.0);
    const double k1 = kC3;
    for (int l = 0; l < nloops; l++) {
        const disco_loopon_loop& lp = in[l];
        if (lp.np < 0) return "negative loopon count";
        if (lp.np > 0 && (!lp.x || !lp.y || !lp.z)) return "null loopon array";
        out.work[l] = (double)lp.np;
        if (lp.np == 0) continue;
        double cx = 0, cy = 0, cz = 0;
        for (int p = 0; p < lp.np; p++) { cx += lp.x[p]; cy += lp.y[p]; cz += lp.z[p]; }
        cx /= lp.np; cy /= lp.np; cz /= lp.np;
        LooponLoop L;
        std::memset(&L, 0, sizeof(L));
        L.cx = cx; L.cy = cy; L.cz = cz;
        const double* ph = lp.phys;
        for (int i = 0; i < 10; i++)
            if (!std::isfinite(ph[i])) return "non-finite loop physics";
        L.bx = ph[0]; L.by = ph[1]; L.bz = ph[2];
        L.mx = ph[3]; L.my = ph[4]; L.mz = ph[5];
        L.nx = ph[6]; L.ny = ph[7]; L.nz = ph[8];
        L.c1 = -k1 * ph[9] * (1.0 - 2.0 * kNu);
        L.c2 = -3.0 * k1 * ph[9];
        L.b_n = ph[3] * ph[6] + ph[4] * ph[7] + ph[5] * ph[8];
        L.p0 = (int)out.pd.size();
        L.np = lp.np;
        double r2max = 0.0;
        for (int p = 0; p < lp.np; p++) {
            const double x = lp.x[p] - cx, y = lp.y[p] - cy, z = lp.z[p] - cz;
            if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z))
                return "non-finite loopon";
            r2max = std::fmax(r2max, x * x + y * y + z * z);
            double4 d; d.x = x; d.y = y; d.z = z; d.w = 0.0;
            float4 f; f.x = (float)x; f.y = (float)y; f.z = (float)z; f.w = 0.0f;
            out.pd.push_back(d);
            out.pf.push_back(f);
        }
        L.rmax2 = (std::sqrt(r2max) + 1e-3) * (std::sqrt(r2max) + 1e-3);
        out.loopons += lp.np;
        out.loops.push_back(L);
    }
    return "";
}

}  // namespace disco
