/*
 * disco_oracle.c -- CPU restatement of DISCO's displacement hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * package's smoke() entry and bench.py's cpu_baseline leg may load it.  The
 * product path (disco_amd, libdisco_hip.so) never calls it and has no CPU
 * fallback.
 *
 * Written from the mathematics of the reference (isotropic Volterra field of a
 * closed polyline, Kroupa loopon term) and its conventions; every function
 * names the reference lines it follows.  Parity status per function is listed
 * in the header of oracle/__init__.py.
 *
 *   oracle_barnett_loop   Code/barnett.hpp:44-107   (FP64, same operation order)
 *   oracle_barnett_loop_f Code/barnett.hpp:44-107   (the same in float, R = float)
 *   oracle_kroupa_pair    Code/kroupa.hpp:52-73     (FP64)
 *   oracle_line_sum       Code/backend_cpu.cpp:38-63  generalised to +-img images
 *   oracle_loopon_sum     Code/backend_cpu.cpp:64-96  generalised to +-img images
 *   oracle_line_sum_ref32 Code/backend_cuda.cu:336-373,476-495 (the reference's
 *                         --fp32 line scheme, restated on the CPU as a yardstick)
 *
 * Images: the reference hard-codes ii, jj, kk in {-1, 0, 1} with kk fastest
 * (Code/backend_cpu.cpp:51-53); here the range is {-img..img} in the same
 * nesting order, so img = 1 is the reference's sum term for term.
 */
#include <math.h>
#include <stddef.h>
#include <omp.h>

#define ORACLE_PI 3.14159265358979323846
/* Poisson ratio, Code/kroupa.hpp:24 */
#define ORACLE_NU 0.35

int oracle_threads(void) { return omp_get_max_threads(); }

/* k1 = 1 / (8 pi (1 - nu)), Code/kroupa.hpp:38-41 */
double oracle_k1(void) { return 1.0 / (8.0 * ORACLE_PI * (1.0 - ORACLE_NU)); }

/*
 * Field at p of the closed polyline V[0..nv-1] (segment k runs V[k-1] -> V[k],
 * V[nv-1] -> V[0] closes it), fan apex c, Burgers vector b; adds to u[0..2].
 *   u += b Omega / (4 pi) + sum_seg [cJ J0 w - cK (d.w)(K0 d + K1 t)]
 * with Omega the fan of Van Oosterom-Strackee triangles (c, A, B) and the
 * cancellation-free branches for J0, K0 (Code/barnett.hpp:64-102).
 */
void oracle_barnett_loop(double px, double py, double pz,
                         const double *vx, const double *vy, const double *vz, int nv,
                         double cx, double cy, double cz,
                         double bx, double by, double bz, double *u)
{
    const double c3 = 1.0 / (8.0 * ORACLE_PI * (1.0 - ORACLE_NU));
    const double cJ = c3 - 1.0 / (4.0 * ORACLE_PI);
    const double cK = c3;

    const double r1x = cx - px, r1y = cy - py, r1z = cz - pz;
    const double l1 = sqrt(r1x * r1x + r1y * r1y + r1z * r1z);

    double omega = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    double ax = vx[nv - 1] - px, ay = vy[nv - 1] - py, az = vz[nv - 1] - pz;
    double la = sqrt(ax * ax + ay * ay + az * az);
    for (int k = 0; k < nv; k++) {
        const double qx = vx[k] - px, qy = vy[k] - py, qz = vz[k] - pz;   /* B - p */
        const double lb = sqrt(qx * qx + qy * qy + qz * qz);

        /* solid angle of triangle (c, A, B) */
        const double num = r1x * (ay * qz - az * qy) + r1y * (az * qx - ax * qz)
                         + r1z * (ax * qy - ay * qx);
        const double den = l1 * la * lb + (r1x * ax + r1y * ay + r1z * az) * lb
                         + (r1x * qx + r1y * qy + r1z * qz) * la
                         + (ax * qx + ay * qy + az * qz) * l1;
        omega += 2.0 * atan2(num, den);

        /* straight segment A -> B */
        double tx = qx - ax, ty = qy - ay, tz = qz - az;
        const double L = sqrt(tx * tx + ty * ty + tz * tz);
        if (L > 0.0) {
            tx /= L; ty /= L; tz /= L;
            const double sA = ax * tx + ay * ty + az * tz;
            const double sB = sA + L;
            const double dx = ax - sA * tx, dy = ay - sA * ty, dz = az - sA * tz;
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double RA = la, RB = lb;
            const double fA = sA >= 0.0 ? RA + sA : d2 / (RA - sA);
            const double fB = sB >= 0.0 ? RB + sB : d2 / (RB - sB);
            const double J0 = log(fB / fA);
            double K0;
            if (sA >= 0.0)      K0 = 1.0 / (RA * (RA + sA)) - 1.0 / (RB * (RB + sB));
            else if (sB <= 0.0) K0 = 1.0 / (RB * (RB - sB)) - 1.0 / (RA * (RA - sA));
            else                K0 = (sB / RB - sA / RA) / d2;
            const double K1 = 1.0 / RA - 1.0 / RB;
            const double wx = by * tz - bz * ty, wy = bz * tx - bx * tz, wz = bx * ty - by * tx;
            const double dw = dx * wx + dy * wy + dz * wz;
            sx += wx * J0 * cJ - cK * dw * (dx * K0 + tx * K1);
            sy += wy * J0 * cJ - cK * dw * (dy * K0 + ty * K1);
            sz += wz * J0 * cJ - cK * dw * (dz * K0 + tz * K1);
        }
        ax = qx; ay = qy; az = qz; la = lb;
    }
    const double fo = omega / (4.0 * ORACLE_PI);
    u[0] += bx * fo + sx;
    u[1] += by * fo + sy;
    u[2] += bz * fo + sz;
}

/* The same evaluation with R = float (Code/barnett.hpp:44-107 instantiated
 * for float, as the reference's --fp32 kernel does): float libm calls. */
void oracle_barnett_loop_f(float px, float py, float pz,
                           const float *vx, const float *vy, const float *vz, int nv,
                           float cx, float cy, float cz,
                           float bx, float by, float bz, float *u)
{
    const float PI = 3.14159265358979323846f;
    const float c3 = (float)(1.0 / (8.0 * ORACLE_PI * (1.0 - ORACLE_NU)));
    const float cJ = c3 - 1.0f / (4.0f * PI);
    const float cK = c3;

    const float r1x = cx - px, r1y = cy - py, r1z = cz - pz;
    const float l1 = sqrtf(r1x * r1x + r1y * r1y + r1z * r1z);
    float omega = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
    float ax = vx[nv - 1] - px, ay = vy[nv - 1] - py, az = vz[nv - 1] - pz;
    float la = sqrtf(ax * ax + ay * ay + az * az);
    for (int k = 0; k < nv; k++) {
        const float qx = vx[k] - px, qy = vy[k] - py, qz = vz[k] - pz;
        const float lb = sqrtf(qx * qx + qy * qy + qz * qz);
        const float num = r1x * (ay * qz - az * qy) + r1y * (az * qx - ax * qz)
                        + r1z * (ax * qy - ay * qx);
        const float den = l1 * la * lb + (r1x * ax + r1y * ay + r1z * az) * lb
                        + (r1x * qx + r1y * qy + r1z * qz) * la
                        + (ax * qx + ay * qy + az * qz) * l1;
        omega += 2.0f * atan2f(num, den);
        float tx = qx - ax, ty = qy - ay, tz = qz - az;
        const float L = sqrtf(tx * tx + ty * ty + tz * tz);
        if (L > 0.0f) {
            tx /= L; ty /= L; tz /= L;
            const float sA = ax * tx + ay * ty + az * tz;
            const float sB = sA + L;
            const float dx = ax - sA * tx, dy = ay - sA * ty, dz = az - sA * tz;
            const float d2 = dx * dx + dy * dy + dz * dz;
            const float RA = la, RB = lb;
            const float fA = sA >= 0.0f ? RA + sA : d2 / (RA - sA);
            const float fB = sB >= 0.0f ? RB + sB : d2 / (RB - sB);
            const float J0 = logf(fB / fA);
            float K0;
            if (sA >= 0.0f)      K0 = 1.0f / (RA * (RA + sA)) - 1.0f / (RB * (RB + sB));
            else if (sB <= 0.0f) K0 = 1.0f / (RB * (RB - sB)) - 1.0f / (RA * (RA - sA));
            else                 K0 = (sB / RB - sA / RA) / d2;
            const float K1 = 1.0f / RA - 1.0f / RB;
            const float wx = by * tz - bz * ty, wy = bz * tx - bx * tz, wz = bx * ty - by * tx;
            const float dw = dx * wx + dy * wy + dz * wz;
            sx += wx * J0 * cJ - cK * dw * (dx * K0 + tx * K1);
            sy += wy * J0 * cJ - cK * dw * (dy * K0 + ty * K1);
            sz += wz * J0 * cJ - cK * dw * (dz * K0 + tz * K1);
        }
        ax = qx; ay = qy; az = qz; la = lb;
    }
    const float fo = omega / (4.0f * PI);
    u[0] += bx * fo + sx;
    u[1] += by * fo + sy;
    u[2] += bz * fo + sz;
}

/*
 * One loopon image at rho = atom - loopon (Code/kroupa.hpp:52-73):
 *   u_i += -k1 dS b_i / rho^2 [ (1-2nu)/rho (n_i m.rho + m_i n.rho - rho_i m.n)
 *                               + 3 rho_i (m.rho)(n.rho) / rho^3 ]
 * phys = {bx, by, bz, mx, my, mz, nx, ny, nz, dS}; rho < 1e-10 adds nothing.
 */
static inline void kroupa_pair_inl(double rx, double ry, double rz, const double *ph,
                                   double k1, double b_n, double *ux, double *uy, double *uz)
{
    const double rho2 = rx * rx + ry * ry + rz * rz;
    const double rho0 = sqrt(rho2);
    const int keep = !(rho0 < 1e-10);
    const double rho = keep ? rho0 : 1.0;
    const double r2 = keep ? rho2 : 1.0;
    const double rho3 = r2 * rho;
    const double m_r = ph[3] * rx + ph[4] * ry + ph[5] * rz;
    const double n_r = ph[6] * rx + ph[7] * ry + ph[8] * rz;
    const double k2 = (1.0 - 2.0 * ORACLE_NU) / rho;
    const double fac = keep ? -k1 * ph[9] / r2 : 0.0;
    const double tail = 3.0 * m_r * n_r / rho3;
    *ux += fac * ph[0] * (k2 * (ph[6] * m_r + ph[3] * n_r - rx * b_n) + tail * rx);
    *uy += fac * ph[1] * (k2 * (ph[7] * m_r + ph[4] * n_r - ry * b_n) + tail * ry);
    *uz += fac * ph[2] * (k2 * (ph[8] * m_r + ph[5] * n_r - rz * b_n) + tail * rz);
}

void oracle_kroupa_pair(double rx, double ry, double rz, const double *phys,
                        double k1, double b_n, double *u)
{
    kroupa_pair_inl(rx, ry, rz, phys, k1, b_n, &u[0], &u[1], &u[2]);
}

/*
 * Line-integral sum (Code/backend_cpu.cpp:38-63): loop-major; for each atom,
 * the images (ii, jj, kk) in [-img, img]^3 with kk fastest, each evaluated as
 * the loop at the atom moved by -delta; summed in registers, then added (+=).
 * Loops are packed: vertices of loop l at voff[l] .. voff[l+1]-1 of vx/vy/vz;
 * c[3l..3l+2] the apex, b[3l..3l+2] the Burgers vector (phys.b * phys.m).
 */
void oracle_line_sum(int n, const double *x, const double *y, const double *z,
                     double *ux, double *uy, double *uz,
                     int nloops, const int *voff, const double *vx, const double *vy,
                     const double *vz, const double *c, const double *b,
                     double Lx, double Ly, double Lz, int img)
{
    for (int l = 0; l < nloops; l++) {
        const int o = voff[l], nv = voff[l + 1] - voff[l];
        if (nv <= 0) continue;
        #pragma omp parallel for schedule(dynamic, 256)
        for (int a = 0; a < n; a++) {
            double u[3] = { 0.0, 0.0, 0.0 };
            for (int ii = -img; ii <= img; ii++)
                for (int jj = -img; jj <= img; jj++)
                    for (int kk = -img; kk <= img; kk++)
                        oracle_barnett_loop(x[a] - Lx * ii, y[a] - Ly * jj, z[a] - Lz * kk,
                                            vx + o, vy + o, vz + o, nv,
                                            c[3 * l], c[3 * l + 1], c[3 * l + 2],
                                            b[3 * l], b[3 * l + 1], b[3 * l + 2], u);
            ux[a] += u[0];
            uy[a] += u[1];
            uz[a] += u[2];
        }
    }
}

/*
 * Loopon sum (Code/backend_cpu.cpp:64-96): loop-major; per atom, images with kk
 * fastest, the loopon loop vectorised with a SIMD reduction.  Loopons of loop l
 * at poff[l] .. poff[l+1]-1 of px/py/pz; phys[10l..10l+9] its LoopPhys.
 */
void oracle_loopon_sum(int n, const double *x, const double *y, const double *z,
                       double *ux, double *uy, double *uz,
                       int nloops, const int *poff, const double *px, const double *py,
                       const double *pz, const double *phys,
                       double Lx, double Ly, double Lz, int img)
{
    const double k1 = oracle_k1();
    for (int l = 0; l < nloops; l++) {
        const int o = poff[l], np = poff[l + 1] - poff[l];
        const double *ph = phys + 10 * l;
        const double b_n = ph[3] * ph[6] + ph[4] * ph[7] + ph[5] * ph[8];
        const double *lx = px + o, *ly = py + o, *lz = pz + o;
        #pragma omp parallel for schedule(dynamic, 256)
        for (int a = 0; a < n; a++) {
            const double ax = x[a], ay = y[a], az = z[a];
            double dux = 0.0, duy = 0.0, duz = 0.0;
            for (int ii = -img; ii <= img; ii++)
                for (int jj = -img; jj <= img; jj++)
                    for (int kk = -img; kk <= img; kk++) {
                        const double ex = Lx * ii, ey = Ly * jj, ez = Lz * kk;
                        #pragma omp simd reduction(+:dux,duy,duz)
                        for (int p = 0; p < np; p++)
                            kroupa_pair_inl(ax - (lx[p] + ex), ay - (ly[p] + ey), az - (lz[p] + ez),
                                            ph, k1, b_n, &dux, &duy, &duz);
                    }
            ux[a] += dux;
            uy[a] += duy;
            uz[a] += duz;
        }
    }
}

/*
 * The reference's single-precision line scheme (Code/backend_cuda.cu:336-373,
 * host packing :476-495), restated on the CPU: vertices as float offsets from
 * the loop centre, the atom reduced to a - c in double then rounded, each
 * image's field in float (R = float), the image fields summed in double.
 * It is the yardstick for "the reference's own FP32 error".
 */
void oracle_line_sum_ref32(int n, const double *x, const double *y, const double *z,
                           double *ux, double *uy, double *uz,
                           int nloops, const int *voff, const double *vx, const double *vy,
                           const double *vz, const double *c, const double *b,
                           double Lx, double Ly, double Lz, int img, float *scratch)
{
    for (int l = 0; l < nloops; l++) {
        const int o = voff[l], nv = voff[l + 1] - voff[l];
        if (nv <= 0) continue;
        float *fx = scratch, *fy = scratch + nv, *fz = scratch + 2 * nv;
        for (int k = 0; k < nv; k++) {
            fx[k] = (float)(vx[o + k] - c[3 * l]);
            fy[k] = (float)(vy[o + k] - c[3 * l + 1]);
            fz[k] = (float)(vz[o + k] - c[3 * l + 2]);
        }
        const float bx = (float)b[3 * l], by = (float)b[3 * l + 1], bz = (float)b[3 * l + 2];
        const float dx = (float)Lx, dy = (float)Ly, dz = (float)Lz;
        #pragma omp parallel for schedule(dynamic, 256)
        for (int a = 0; a < n; a++) {
            const float qx = (float)(x[a] - c[3 * l]);
            const float qy = (float)(y[a] - c[3 * l + 1]);
            const float qz = (float)(z[a] - c[3 * l + 2]);
            double sx = 0.0, sy = 0.0, sz = 0.0;
            for (int ii = -img; ii <= img; ii++)
                for (int jj = -img; jj <= img; jj++)
                    for (int kk = -img; kk <= img; kk++) {
                        float f[3] = { 0.0f, 0.0f, 0.0f };
                        oracle_barnett_loop_f(qx - dx * (float)ii, qy - dy * (float)jj,
                                              qz - dz * (float)kk, fx, fy, fz, nv,
                                              0.0f, 0.0f, 0.0f, bx, by, bz, f);
                        sx += (double)f[0];
                        sy += (double)f[1];
                        sz += (double)f[2];
                    }
            ux[a] += sx;
            uy[a] += sy;
            uz[a] += sz;
        }
    }
}
