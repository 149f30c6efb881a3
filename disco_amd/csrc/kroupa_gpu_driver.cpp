// kroupa_gpu_driver.cpp -- a caller of the drop-in GPU backend shaped like the
// reference driver's backend call (Code/main.cpp:689-730): it assembles
// AtomArrays and std::vector<LoopData> (Code/backend.hpp types), zeroes the
// displacements, checks gpuAvailable(), prints the banner and calls kroupaGPU.
//
//   kroupa_gpu_driver <input.bin> <output.bin>
//
// input (little endian): int32 n; f64 x[n], y[n], z[n]; f64 box[3];
//   int32 fp32; int32 nloops; per loop: int32 line; f64 phys[10];
//   line: int32 nv; f64 vx[nv], vy[nv], vz[nv]; f64 c[3]
//   loopons: int32 np; f64 x[np], y[np], z[np]
// output: f64 ux[n], uy[n], uz[n]; f64 loopSec[nloops]; f64 upload, kernels, download
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "backend.hpp"

using namespace std;

namespace {
FILE* in = nullptr;
template <typename T> T rd()
{
    T v;
    if (fread(&v, sizeof(T), 1, in) != 1) { cerr << "driver: truncated input\n"; exit(2); }
    return v;
}
vector<double> rdv(size_t n)
{
    vector<double> v(n);
    if (n && fread(v.data(), sizeof(double), n, in) != n) { cerr << "driver: truncated input\n"; exit(2); }
    return v;
}
}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) { cerr << "usage: " << argv[0] << " <input.bin> <output.bin>\n"; return 1; }
    in = fopen(argv[1], "rb");
    if (!in) { cerr << "driver: cannot open " << argv[1] << "\n"; return 1; }
    const int n = rd<int>();
    vector<double> x = rdv(n), y = rdv(n), z = rdv(n);
    const vector<double> box = rdv(3);
    GpuOptions opt;
    opt.fp32 = rd<int>() != 0;
    const int nloops = rd<int>();
    vector<LoopData> loops(nloops);
    for (int l = 0; l < nloops; l++) {
        LoopData& d = loops[l];
        d.line = rd<int>() != 0;
        const vector<double> ph = rdv(10);
        d.phys = { ph[0], ph[1], ph[2], ph[3], ph[4], ph[5], ph[6], ph[7], ph[8], ph[9] };
        const int m = rd<int>();
        if (d.line) {
            d.vx = rdv(m); d.vy = rdv(m); d.vz = rdv(m);
            const vector<double> c = rdv(3);
            d.cx = c[0]; d.cy = c[1]; d.cz = c[2];
        } else {
            d.x = rdv(m); d.y = rdv(m); d.z = rdv(m);
        }
    }
    fclose(in);

    if (!gpuAvailable()) { cerr << "driver: no usable GPU\n"; return 3; }
    printGpuInfo();
    vector<double> ux(n, 0.), uy(n, 0.), uz(n, 0.);
    const AtomArrays atoms = { x.data(), y.data(), z.data(), ux.data(), uy.data(), uz.data(), n };
    BackendTiming t;
    kroupaGPU(atoms, loops, box[0], box[1], box[2], opt, t);
    if ((int)t.loopSec.size() != nloops) { cerr << "driver: loopSec has the wrong size\n"; return 4; }

    FILE* out = fopen(argv[2], "wb");
    if (!out) { cerr << "driver: cannot open " << argv[2] << "\n"; return 1; }
    fwrite(ux.data(), sizeof(double), n, out);
    fwrite(uy.data(), sizeof(double), n, out);
    fwrite(uz.data(), sizeof(double), n, out);
    fwrite(t.loopSec.data(), sizeof(double), nloops, out);
    const double tt[3] = { t.upload, t.kernels, t.download };
    fwrite(tt, sizeof(double), 3, out);
    fclose(out);
    return 0;
}
