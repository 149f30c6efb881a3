// backend_hip.cpp -- DISCO's GPU backend entry points for MI355X, on the C ABI
// of include/disco_hip.h.
//
// Defines, with C++ linkage and exactly the signatures of Code/backend.hpp:57-63
// (compiled with -DWITH_CUDA, the macro the unmodified Code/main.cpp tests):
//   bool gpuAvailable();     replaces Code/backend_cuda.cu:537-541
//   void printGpuInfo();     replaces Code/backend_cuda.cu:543-553 (CUs, not SMs)
//   bool gpuLineFp64();      replaces Code/backend_cuda.cu:525-532 (gfx950 has FP64)
//   void kroupaGPU(...);     replaces Code/backend_cuda.cu:568-722
// kroupaCPU / cpuThreads stay in the reference's Code/backend_cpu.cpp.
//
// It is compiled against the reference's own Code/backend.hpp (include path),
// so a type mismatch with the interface is a compile error here.
//
// Behaviour kept from kroupaGPU: the same stdout lines ("Precision: ...",
// "Running GPU kernels...", "Loop l/N: ..." per loop, "GPU kernel time"), the
// BackendTiming fields (upload, kernels, download, loopSec of size nLoops),
// and print-and-exit on any device error (Code/backend_cuda.cu:45-53).  The
// displacement arrays end holding the sum (the caller zeroes them,
// Code/main.cpp:714).  GpuOptions::fp32 selects DISCO_PREC_FP32C;
// GpuOptions::chunkPairs is not needed (each loop is one launch over all
// atoms, and there is no watchdog) and is reported when set.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "backend.hpp"
#include "disco_hip.h"

using namespace std;

namespace {

void die(int code, disco_ctx* ctx, const char* what, int line)
{
    cerr << "HIP backend error at " << __FILE__ << ":" << line << "  " << what << ": "
         << disco_hip_strerror(code);
    if (ctx && disco_hip_last_error(ctx)[0]) cerr << " (" << disco_hip_last_error(ctx) << ")";
    cerr << "\n";
    exit(EXIT_FAILURE);
}

#define DISCO_CHECK(ctx, call)                              \
    do {                                                    \
        const int _rc = (call);                             \
        if (_rc != DISCO_OK) die(_rc, (ctx), #call, __LINE__); \
    } while (0)

}  // namespace

bool gpuAvailable()
{
    int n = 0;
    return disco_hip_device_count(&n) == DISCO_OK && n > 0;
}

void printGpuInfo()
{
    disco_device_info info;
    DISCO_CHECK(nullptr, disco_hip_device_info(0, &info));
    cout << "\n--- GPU: " << info.name << " (" << info.arch << ") ---\n";
    cout << "  CUs              : " << info.compute_units << "\n";
    cout << "  Global memory    : " << info.global_mem / (1 << 20) << " MB\n";
    cout << "  LDS/block        : " << info.lds_per_block / 1024 << " KB\n\n";
}

bool gpuLineFp64() { return disco_hip_line_fp64() != 0; }

void kroupaGPU(const AtomArrays& atoms, const vector<LoopData>& loops, double dx, double dy,
               double dz, const GpuOptions& opt, BackendTiming& t)
{
    const int nbLoops = (int)loops.size();
    const int prec = opt.fp32 ? DISCO_PREC_FP32C : DISCO_PREC_FP64;
    cout << "Precision: "
         << (opt.fp32 ? "single (fp32 pair terms, fp64 image sums, fp64 near the cut)" : "double")
         << "\n";
    if (opt.chunkPairs > 0.)
        cout << "Chunking : not needed on this backend (one launch per loop); ignored\n";

    disco_ctx* ctx = nullptr;
    DISCO_CHECK(nullptr, disco_hip_open(0, &ctx));

    // Split by method, keeping each loop's index for loopSec
    vector<disco_line_loop> line;
    vector<disco_loopon_loop> lps;
    vector<int> lineIdx, lpIdx;
    for (int l = 0; l < nbLoops; l++) {
        const LoopData& lp = loops[l];
        if (lp.line) {
            disco_line_loop d;
            d.nv = (int)lp.vx.size();
            d.vx = lp.vx.data(); d.vy = lp.vy.data(); d.vz = lp.vz.data();
            d.c[0] = lp.cx; d.c[1] = lp.cy; d.c[2] = lp.cz;
            d.b[0] = lp.phys.bx * lp.phys.mx;
            d.b[1] = lp.phys.by * lp.phys.my;
            d.b[2] = lp.phys.bz * lp.phys.mz;
            line.push_back(d);
            lineIdx.push_back(l);
        } else {
            disco_loopon_loop d;
            d.np = (int)lp.x.size();
            d.x = lp.x.data(); d.y = lp.y.data(); d.z = lp.z.data();
            const LoopPhys& P = lp.phys;
            const double ph[10] = { P.bx, P.by, P.bz, P.mx, P.my, P.mz, P.nx, P.ny, P.nz, P.dS };
            for (int i = 0; i < 10; i++) d.phys[i] = ph[i];
            lps.push_back(d);
            lpIdx.push_back(l);
        }
    }

    cout << "Running GPU kernels...\n";
    t.loopSec.assign(nbLoops, 0.);
    t.upload = t.kernels = t.download = 0.;
    for (int pass = 0; pass < 2; pass++) {
        const vector<int>& idx = pass == 0 ? lineIdx : lpIdx;
        if (idx.empty()) continue;
        vector<double> sec(idx.size(), 0.);
        disco_timing tm = { 0., 0., 0., 0., sec.data() };
        if (pass == 0)
            DISCO_CHECK(ctx, disco_hip_line_sum(ctx, atoms.n, atoms.x, atoms.y, atoms.z, atoms.ux,
                                                atoms.uy, atoms.uz, line.data(), (int)line.size(),
                                                dx, dy, dz, 1, prec, &tm));
        else
            DISCO_CHECK(ctx, disco_hip_loopon_sum(ctx, atoms.n, atoms.x, atoms.y, atoms.z,
                                                  atoms.ux, atoms.uy, atoms.uz, lps.data(),
                                                  (int)lps.size(), dx, dy, dz, 1, prec, &tm));
        t.upload += tm.upload;
        t.kernels += tm.kernels;
        t.download += tm.download;
        for (size_t k = 0; k < idx.size(); k++) {
            const int l = idx[k];
            t.loopSec[l] = sec[k];
            const LoopData& lp = loops[l];
            cout << "  Loop " << l + 1 << "/" << nbLoops << ":  "
                 << (lp.line ? lp.vx.size() : lp.x.size())
                 << (lp.line ? " segments" : " loopons") << "  " << t.loopSec[l] << " s\n";
        }
    }
    cout << "\nGPU kernel time: " << t.kernels * 1e3 << " ms\n\n";
    DISCO_CHECK(nullptr, disco_hip_close(ctx));
}
