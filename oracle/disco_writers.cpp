/*
 * disco_writers.cpp -- CPU restatement of DISCO's output stage.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker of the device writers:
 * only tests/ and bench.py's cpu_baseline leg may load it.  The product path
 * (disco_amd, libdisco_hip.so) never calls it and has no CPU fallback.
 *
 *   oracle_final_positions  Code/main.cpp:797-818  (displacement, wrap into the
 *                           box centred at the origin, shift by +L/2, apply U)
 *   oracle_write_xyz        Code/main.cpp:820-845  (extended XYZ)
 *   oracle_write_lammps     Code/main.cpp:50-135   (writeLAMMPS, atomic style)
 *   oracle_lammps_box       Code/main.cpp:64-86    (Gram-Schmidt box parameters)
 *
 * The text goes through an iostream with the reference's manipulators (the
 * format IS the specification), and the arithmetic keeps the reference's
 * expression shapes.  Built like the reference's host code: g++ -O3 with FMA
 * (x86-64-v3) and GCC's default contraction, which fuses a*b + c*d + e*f into
 * fma(e, f, fma(a, b, c*d)) and a - b*c into fma(-b, c, a); the tests pin
 * those results against exact rational arithmetic.
 */
#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>

namespace {

struct Frame {
    double lx, ly, lz, xy, xz, yz;
    double e1[3], e2[3], e3[3];
};

// Box vectors are the columns of Hp (row-major 3x3).
Frame gram_schmidt(const double* Hp)
{
    Frame f;
    const double a1[3] = {Hp[0], Hp[3], Hp[6]};
    const double a2[3] = {Hp[1], Hp[4], Hp[7]};
    const double a3[3] = {Hp[2], Hp[5], Hp[8]};
    f.lx = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
    for (int i = 0; i < 3; i++) f.e1[i] = a1[i] / f.lx;
    f.xy = a2[0] * f.e1[0] + a2[1] * f.e1[1] + a2[2] * f.e1[2];
    double r2[3];
    for (int i = 0; i < 3; i++) r2[i] = a2[i] - f.xy * f.e1[i];
    f.ly = sqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
    for (int i = 0; i < 3; i++) f.e2[i] = r2[i] / f.ly;
    f.xz = a3[0] * f.e1[0] + a3[1] * f.e1[1] + a3[2] * f.e1[2];
    double r3[3];
    for (int i = 0; i < 3; i++) r3[i] = a3[i] - f.xz * f.e1[i];
    f.yz = r3[0] * f.e2[0] + r3[1] * f.e2[1] + r3[2] * f.e2[2];
    double r3p[3];
    for (int i = 0; i < 3; i++) r3p[i] = r3[i] - f.yz * f.e2[i];
    f.lz = sqrt(r3p[0] * r3p[0] + r3p[1] * r3p[1] + r3p[2] * r3p[2]);
    for (int i = 0; i < 3; i++) f.e3[i] = r3p[i] / f.lz;
    return f;
}

}  // namespace

extern "C" {

// U row-major; box = (dx, dy, dz).
void oracle_final_positions(long long n, const double* x, const double* y, const double* z,
                            const double* ux, const double* uy, const double* uz,
                            const double* box, const double* U, double* fx, double* fy,
                            double* fz)
{
    const double dx = box[0], dy = box[1], dz = box[2];
    const double hx = dx / 2., hy = dy / 2., hz = dz / 2.;
    for (long long k = 0; k < n; k++) {
        double px = x[k] + ux[k];
        double py = y[k] + uy[k];
        double pz = z[k] + uz[k];
        if (px > hx) px -= dx; else if (px < -hx) px += dx;
        if (py > hy) py -= dy; else if (py < -hy) py += dy;
        if (pz > hz) pz -= dz; else if (pz < -hz) pz += dz;
        const double sx = px + hx, sy = py + hy, sz = pz + hz;
        fx[k] = U[0] * sx + U[1] * sy + U[2] * sz;
        fy[k] = U[3] * sx + U[4] * sy + U[5] * sz;
        fz[k] = U[6] * sx + U[7] * sy + U[8] * sz;
    }
}

// out6 = (lx, ly, lz, xy, xz, yz), frame9 = (e1, e2, e3) rows.
void oracle_lammps_box(const double* Hp, double* out6, double* frame9)
{
    const Frame f = gram_schmidt(Hp);
    const double v[6] = {f.lx, f.ly, f.lz, f.xy, f.xz, f.yz};
    for (int i = 0; i < 6; i++) out6[i] = v[i];
    for (int i = 0; i < 3; i++) {
        frame9[i] = f.e1[i];
        frame9[3 + i] = f.e2[i];
        frame9[6 + i] = f.e3[i];
    }
}

// grain / twin may be NULL (no column).  Returns 0, or 1 if the file cannot
// be opened.
int oracle_write_xyz(const char* path, long long n, const double* fx, const double* fy,
                     const double* fz, const double* Hp, const char* element, const int* grain,
                     const int* twin)
{
    std::ofstream out(path);
    if (!out) return 1;
    out << n << "\n";
    out << "Lattice=\"";
    for (int c = 0; c < 3; c++) {
        if (c) out << "  ";
        out << Hp[c] << " " << Hp[3 + c] << " " << Hp[6 + c];
    }
    out << "\" Properties=species:S:1:pos:R:3" << (grain ? ":grain:I:1" : "")
        << (twin ? ":twin:I:1" : "") << "\n";
    out << std::fixed << std::setprecision(8);
    for (long long k = 0; k < n; k++) {
        out << element << " " << fx[k] << " " << fy[k] << " " << fz[k];
        if (grain) out << " " << grain[k] + 1;
        if (twin) out << " " << twin[k];
        out << "\n";
    }
    return out.good() ? 0 : 1;
}

int oracle_write_lammps(const char* path, long long n, const double* fx, const double* fy,
                        const double* fz, const double* Hp, const char* element, double mass)
{
    const Frame f = gram_schmidt(Hp);
    std::ofstream out(path);
    if (!out) return 1;
    out << std::fixed << std::setprecision(8);
    out << "# LAMMPS data file generated by exec_cuda\n\n";
    out << n << " atoms\n\n";
    out << "1 atom types\n\n";
    out << "0.0 " << f.lx << " xlo xhi\n";
    out << "0.0 " << f.ly << " ylo yhi\n";
    out << "0.0 " << f.lz << " zlo zhi\n";
    const double tol = 1e-8;
    if (fabs(f.xy) > tol || fabs(f.xz) > tol || fabs(f.yz) > tol)
        out << f.xy << " " << f.xz << " " << f.yz << " xy xz yz\n";
    out << "\nMasses\n\n";
    out << "1 " << std::defaultfloat << std::setprecision(10) << mass << std::fixed
        << std::setprecision(8) << "  # " << element << "\n";
    out << "\nAtoms  # atomic\n\n";
    for (long long k = 0; k < n; k++) {
        const double x = fx[k], y = fy[k], z = fz[k];
        out << k + 1 << " 1 " << f.e1[0] * x + f.e1[1] * y + f.e1[2] * z << " "
            << f.e2[0] * x + f.e2[1] * y + f.e2[2] * z << " "
            << f.e3[0] * x + f.e3[1] * y + f.e3[2] * z << "\n";
    }
    return out.good() ? 0 : 1;
}

// One value as the writers print it (fixed, 8 decimals) into buf; returns its length.
int oracle_fixed8(double v, char* buf, int cap)
{
    std::ostringstream s;
    s << std::fixed << std::setprecision(8) << v;
    const std::string t = s.str();
    if ((int)t.size() >= cap) return -1;
    for (size_t i = 0; i < t.size(); i++) buf[i] = t[i];
    buf[t.size()] = 0;
    return (int)t.size();
}

}  // extern "C"
