// prismatic.hip -- prismatic slab selection on gfx950 (include/disco_prismatic.h).
//
// Reference behaviour replaced: prismaticSlab (Code/prismatic.cpp:30-74), an
// O(atoms) host scan per loop with minImage, the signed distance to the cut
// and loop::contains / insideUV (Code/loop.cpp:281-330).
//
// Three launches per loop, no waits between workgroups:
//   slab_flag_kernel     a workgroup per tile of 4096 atoms, one atom per lane
//                        per step: the predicate, a flag byte per atom
//                        (0 / +1 / -1) and the tile's count (wave ballots);
//   slab_scan_kernel     exclusive scan of the tile counts (one workgroup,
//                        ~1000 counts for 4M atoms);
//   slab_scatter_kernel  each selected atom writes its index and side at its
//                        tile offset + its rank in the tile: ascending order,
//                        as the reference's push_back.
// HBM traffic per atom: 24 B of positions + 1 B flag written, then 1 B read.
//
// Rounding: the predicate is the reference's expression by expression; the
// products its host build fuses into FMAs (GCC, x86-64 with FMA) are fused
// here with fma() in the same places and the rest is not contracted, and
// double division / sqrt are correctly rounded on this device, so every
// comparison sees the reference's values bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "ctx_internal.h"
#include "disco_prismatic.h"

#pragma clang fp contract(off)

using namespace disco;

namespace {

constexpr int kSlabBlock = 256;
constexpr int kSlabTile = 4096;   // atoms per workgroup (16 per lane)
constexpr int kScanThreads = 1024;

// The loop as the predicate reads it (device copy).
struct SlabLoop {
    disco_loop_state s;
    double cuv;       // u . v
    double det;       // 1 - (u . v)^2
    double t_lo, t_hi;
    double spacing;
    double lx, ly, lz;
};

__device__ __forceinline__ double dmax(double a, double b) { return a < b ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }

// r - L floor(r / L + 1/2)  (Code/prismatic.cpp:28-31), the product fused
__device__ __forceinline__ double min_image(double r, double L)
{
    return fma(-L, floor(r / L + 0.5), r);
}

// Code/loop.cpp:281-316
__device__ bool inside_uv(const disco_loop_state& L, double u, double v)
{
    const double eps = 1e-9;
    if (L.shape == DISCO_SHAPE_CIRCLE) return sqrt(fma(u, u, v * v)) < L.radius - eps;
    if (L.shape == DISCO_SHAPE_POLYGON) {
        const int N = L.nvert;
        bool in_q = true;
        for (int i = 0; i < N && in_q; i++)
            if (fma(L.nu[i], u - L.pu[i], L.nv[i] * (v - L.pv[i])) < L.r_corner - eps) in_q = false;
        if (in_q) return true;
        if (L.r_corner <= 0.0) return false;
        double d2 = 1e300;
        for (int i = 0; i < N; i++) {
            const int j = i + 1 == N ? 0 : i + 1;
            const double eu = L.qu[j] - L.qu[i], ev = L.qv[j] - L.qv[i];
            const double l2 = fma(eu, eu, ev * ev);
            double t = l2 > 0.0 ? fma(u - L.qu[i], eu, (v - L.qv[i]) * ev) / l2 : 0.0;
            t = dmin(dmax(t, 0.0), 1.0);
            const double wu = fma(-t, eu, u - L.qu[i]);
            const double wv = fma(-t, ev, v - L.qv[i]);
            d2 = dmin(d2, fma(wu, wu, wv * wv));
        }
        return d2 <= fma(L.r_corner, L.r_corner, eps);
    }
    const double du = dmax(fabs(u) - dmax(L.half_m - L.r_corner, 0.0), 0.0);
    const double dv = dmax(fabs(v) - dmax(L.half_p - L.r_corner, 0.0), 0.0);
    return fma(du, du, dv * dv) <= fma(L.r_corner, L.r_corner, eps);
}

// +1 / -1 for a slab atom (side of the cut), 0 otherwise
__device__ __forceinline__ int slab_side(const SlabLoop& P, double x, double y, double z)
{
    const disco_loop_state& L = P.s;
    const double r0 = min_image(x - L.center[0], P.lx);
    const double r1 = min_image(y - L.center[1], P.ly);
    const double r2 = min_image(z - L.center[2], P.lz);
    const double s = fma(r2, L.n[2], fma(r0, L.n[0], r1 * L.n[1]));
    const double t = s / P.spacing;
    if (t < P.t_lo || t >= P.t_hi) return 0;
    // loop::contains: oblique coordinates of rel on (u, v)
    const double ru = fma(r2, L.u[2], fma(r0, L.u[0], r1 * L.u[1]));
    const double rv = fma(r2, L.v[2], fma(r0, L.v[0], r1 * L.v[1]));
    if (!inside_uv(L, fma(-P.cuv, rv, ru) / P.det, fma(-P.cuv, ru, rv) / P.det)) return 0;
    return s > 0.0 ? 1 : -1;
}

// Block-wide exclusive rank of this lane's flag among the workgroup's set
// flags (wave ballots, wave counts in LDS); returns the workgroup's total.
__device__ __forceinline__ unsigned block_rank(bool set, unsigned* wcount, unsigned* rank)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long ball = __ballot(set);
    if (lane == 0) wcount[wave] = (unsigned)__popcll(ball);
    __syncthreads();
    unsigned before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kSlabBlock / 64; w++) {
        before += w < wave ? wcount[w] : 0u;
        total += wcount[w];
    }
    *rank = before + (unsigned)__popcll(ball & ((1ull << lane) - 1ull));
    __syncthreads();   // wcount is rewritten by the next call
    return total;
}

// One workgroup per tile of kSlabTile atoms (kSlabBlock lanes, kSlabTile /
// kSlabBlock atoms each): the flag of every atom, and the tile's slab count.
__global__ __launch_bounds__(kSlabBlock) void slab_flag_kernel(
    const SlabLoop* __restrict__ loop, long long n, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, signed char* __restrict__ flags,
    unsigned* __restrict__ tile_counts)
{
    __shared__ unsigned wcount[kSlabBlock / 64];
    const long long t0 = (long long)blockIdx.x * kSlabTile;
    unsigned count = 0;
    for (int j = 0; j < kSlabTile; j += kSlabBlock) {
        const long long k = t0 + j + threadIdx.x;
        int f = 0;
        if (k < n) {
            f = slab_side(*loop, x[k], y[k], z[k]);
            flags[k] = (signed char)f;
        }
        count += (unsigned)__popcll(__ballot(f != 0));   // wave-uniform
    }
    if ((threadIdx.x & 63) == 0) wcount[threadIdx.x >> 6] = count;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0;
        for (int w = 0; w < kSlabBlock / 64; w++) t += wcount[w];
        tile_counts[blockIdx.x] = t;
    }
}

// Exclusive scan of the tile counts, one workgroup, one count per lane per
// round (coalesced), a running carry between rounds.
__global__ __launch_bounds__(kScanThreads) void slab_scan_kernel(
    const unsigned* __restrict__ counts, unsigned long long* __restrict__ offsets, long long nb,
    unsigned long long* __restrict__ total)
{
    __shared__ unsigned long long wsum[kScanThreads / 64];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    unsigned long long carry = 0;
    for (long long base = 0; base < nb; base += kScanThreads) {
        const long long i = base + t;
        const unsigned long long mine = i < nb ? counts[i] : 0ull;
        unsigned long long inc = mine;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned long long v = __shfl_up(inc, o);
            if (lane >= o) inc += v;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        unsigned long long before = 0, round = 0;
#pragma unroll
        for (int w = 0; w < kScanThreads / 64; w++) {
            before += w < wave ? wsum[w] : 0ull;
            round += wsum[w];
        }
        if (i < nb) offsets[i] = carry + before + inc - mine;
        carry += round;
        __syncthreads();   // wsum is rewritten next round
    }
    if (t == 0) *total = carry;
}

// Each selected atom writes its index and side at its tile's offset plus its
// rank in the tile: ascending order, as the reference's push_back.
__global__ __launch_bounds__(kSlabBlock) void slab_scatter_kernel(
    long long n, const signed char* __restrict__ flags,
    const unsigned long long* __restrict__ offsets, int* __restrict__ atoms,
    int* __restrict__ side, long long cap)
{
    __shared__ unsigned wcount[kSlabBlock / 64];
    const long long t0 = (long long)blockIdx.x * kSlabTile;
    unsigned long long pos0 = offsets[blockIdx.x];
    for (int j = 0; j < kSlabTile; j += kSlabBlock) {
        const long long k = t0 + j + threadIdx.x;
        const int f = k < n ? flags[k] : 0;
        unsigned rank;
        const unsigned step = block_rank(f != 0, wcount, &rank);
        if (f) {
            const unsigned long long pos = pos0 + rank;
            if ((long long)pos < cap) {
                atoms[pos] = (int)k;
                side[pos] = f;
            }
        }
        pos0 += step;
    }
}

// ---------------------------------------------------------------------------
// Applying the slabs (Code/main.cpp:586-687, 733-742): order-keeping
// compactions built from the same three steps -- per-tile counts of a
// predicate, the scan of the counts, a scatter at tile offset + rank.  Pred is
// a functor (bool)(k); Move (k, keep, pos) writes item k's data at pos.

template <class Pred>
__global__ __launch_bounds__(kSlabBlock) void keep_count_kernel(Pred pred, long long n,
                                                                unsigned* __restrict__ tile_counts)
{
    __shared__ unsigned wcount[kSlabBlock / 64];
    const long long t0 = (long long)blockIdx.x * kSlabTile;
    unsigned count = 0;
    for (int j = 0; j < kSlabTile; j += kSlabBlock) {
        const long long k = t0 + j + threadIdx.x;
        count += (unsigned)__popcll(__ballot(k < n && pred(k)));
    }
    if ((threadIdx.x & 63) == 0) wcount[threadIdx.x >> 6] = count;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0;
        for (int w = 0; w < kSlabBlock / 64; w++) t += wcount[w];
        tile_counts[blockIdx.x] = t;
    }
}

template <class Pred, class Move>
__global__ __launch_bounds__(kSlabBlock) void keep_scatter_kernel(
    Pred pred, Move move, long long n, const unsigned long long* __restrict__ offsets)
{
    __shared__ unsigned wcount[kSlabBlock / 64];
    const long long t0 = (long long)blockIdx.x * kSlabTile;
    unsigned long long pos0 = offsets[blockIdx.x];
    for (int j = 0; j < kSlabTile; j += kSlabBlock) {
        const long long k = t0 + j + threadIdx.x;
        const bool keep = k < n && pred(k);
        unsigned rank;
        const unsigned step = block_rank(keep, wcount, &rank);
        if (k < n) move(k, keep, pos0 + rank);
        pos0 += step;
    }
}

// slab entries of one grain and twin region (Code/main.cpp:605-615)
struct KeepRegion {
    const int* atoms;
    const int* grain;   // null: any grain
    const int* twin;    // null: any region
    int g, region;
    long long natoms;   // entries naming no atom are dropped
    __device__ bool operator()(long long q) const
    {
        const int a = atoms[q];
        if (a < 0 || a >= natoms) return false;
        return (!grain || grain[a] == g) && (!twin || twin[a] == region);
    }
};
struct MoveSlab {
    const int *atoms, *side;
    int *atoms_o, *side_o;
    __device__ void operator()(long long q, bool keep, unsigned long long pos) const
    {
        if (keep) { atoms_o[pos] = atoms[q]; side_o[pos] = side[q]; }
    }
};

// atoms not removed (Code/main.cpp:664-676), with the new index of every atom
struct KeepAtom {
    const unsigned char* removed;
    __device__ bool operator()(long long a) const { return removed[a] == 0; }
};
struct MoveAtom {
    const double *x, *y, *z;
    const int* label;
    double *xo, *yo, *zo;
    int* label_o;
    int* new_index;
    __device__ void operator()(long long a, bool keep, unsigned long long pos) const
    {
        new_index[a] = keep ? (int)pos : -1;
        if (!keep) return;
        xo[pos] = x[a];
        yo[pos] = y[a];
        zo[pos] = z[a];
        if (label) label_o[pos] = label[a];
    }
};

// copies whose source survives, renumbered (Code/main.cpp:679-682)
struct KeepCopy {
    const int* copy_atom;
    const int* new_index;
    long long n;   // copies naming no atom are dropped
    __device__ bool operator()(long long k) const
    {
        const int a = copy_atom[k];
        return a >= 0 && a < n && new_index[a] >= 0;
    }
};
struct MoveCopy {
    const int* copy_atom;
    const int* new_index;
    const double* shift;
    int* atom_o;
    double* shift_o;
    __device__ void operator()(long long k, bool keep, unsigned long long pos) const
    {
        if (!keep) return;
        atom_o[pos] = new_index[copy_atom[k]];
        shift_o[3 * pos] = shift[3 * k];
        shift_o[3 * pos + 1] = shift[3 * k + 1];
        shift_o[3 * pos + 2] = shift[3 * k + 2];
    }
};

// removed[atoms[q]] = 1 for a vacancy slab (Code/main.cpp:626-628)
__global__ void mark_removed_kernel(long long natoms, long long count,
                                    const int* __restrict__ atoms,
                                    unsigned char* __restrict__ removed)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const int a = atoms[q];
    if (a >= 0 && a < natoms) removed[a] = 1;
}

// the copies of an interstitial slab: source and shift by side
// (Code/main.cpp:652-655)
__global__ void slab_copies_kernel(long long count, const int* __restrict__ atoms,
                                   const int* __restrict__ side, double p0, double p1, double p2,
                                   double m0, double m1, double m2, int* __restrict__ copy_atom,
                                   double* __restrict__ copy_shift)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const bool plus = side[q] > 0;
    copy_atom[q] = atoms[q];
    copy_shift[3 * q] = plus ? p0 : m0;
    copy_shift[3 * q + 1] = plus ? p1 : m1;
    copy_shift[3 * q + 2] = plus ? p2 : m2;
}

// entry n + k = copy k: the source's site and label, its displacement plus the
// shift (Code/main.cpp:733-742)
__global__ void append_copies_kernel(long long n, long long ncopies,
                                     const int* __restrict__ copy_atom,
                                     const double* __restrict__ shift, double* __restrict__ x,
                                     double* __restrict__ y, double* __restrict__ z,
                                     double* __restrict__ ux, double* __restrict__ uy,
                                     double* __restrict__ uz, int* __restrict__ label)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncopies) return;
    const int a = copy_atom[k];
    if (a < 0 || a >= n) return;   // names no atom: entry n + k is not written
    const long long o = n + k;
    x[o] = x[a];
    y[o] = y[a];
    z[o] = z[a];
    ux[o] = ux[a] + shift[3 * k];
    uy[o] = uy[a] + shift[3 * k + 1];
    uz[o] = uz[a] + shift[3 * k + 2];
    if (label) label[o] = label[a];
}

int ensure(disco_ctx* ctx, void** p, size_t* cap, size_t bytes)
{
    if (*cap >= bytes) return DISCO_OK;
    if (*p) DISCO_HIP_TRY(ctx, hipFree(*p));
    *p = nullptr;
    *cap = 0;
    DISCO_HIP_TRY(ctx, hipMalloc(p, bytes));
    *cap = bytes;
    return DISCO_OK;
}

// Scratch of a compaction over at most nmax items, plus `extra` bytes after it.
struct Compaction {
    unsigned long long* total;
    unsigned* counts;
    unsigned long long* offs;
    char* extra;
};

int compaction_scratch(disco_ctx* ctx, long long nmax, size_t extra, Compaction* c)
{
    const long long nb = (nmax + kSlabTile - 1) / kSlabTile + 1;
    const size_t o_counts = 256;
    const size_t o_offs = (o_counts + (size_t)nb * 4 + 255) & ~(size_t)255;
    const size_t o_extra = (o_offs + (size_t)nb * 8 + 255) & ~(size_t)255;
    const int rc = ensure(ctx, &ctx->d_slab_scratch, &ctx->cap_slab_scratch, o_extra + extra);
    if (rc) return rc;
    char* sc = static_cast<char*>(ctx->d_slab_scratch);
    c->total = reinterpret_cast<unsigned long long*>(sc);
    c->counts = reinterpret_cast<unsigned*>(sc + o_counts);
    c->offs = reinterpret_cast<unsigned long long*>(sc + o_offs);
    c->extra = sc + o_extra;
    return DISCO_OK;
}

// The pinned word the compaction totals come back through.
int pinned_total(disco_ctx* ctx, unsigned long long** h)
{
    if (!ctx->h_slab_stage)
        DISCO_HIP_TRY(ctx, hipHostMalloc(&ctx->h_slab_stage, sizeof(SlabLoop) + 256,
                                         hipHostMallocDefault));
    *h = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->h_slab_stage) +
                                               ((sizeof(SlabLoop) + 63) & ~(size_t)63));
    return DISCO_OK;
}

// Items k < n for which pred(k) holds, moved in order; the count is left in
// c.total on the device (stream-ordered).
template <class Pred, class Move>
int launch_compaction(disco_ctx* ctx, long long n, Pred pred, Move move, const Compaction& c,
                      hipStream_t s)
{
    const long long nb = (n + kSlabTile - 1) / kSlabTile;
    if (nb > 0) {
        hipLaunchKernelGGL((keep_count_kernel<Pred>), dim3((unsigned)nb), dim3(kSlabBlock), 0, s,
                           pred, n, c.counts);
        DISCO_HIP_TRY(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(slab_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, c.counts, c.offs, nb,
                       c.total);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    if (nb > 0) {
        hipLaunchKernelGGL((keep_scatter_kernel<Pred, Move>), dim3((unsigned)nb),
                           dim3(kSlabBlock), 0, s, pred, move, n, c.offs);
        DISCO_HIP_TRY(ctx, hipGetLastError());
    }
    return DISCO_OK;
}

unsigned blocks256(long long n) { return (unsigned)((n + 255) / 256); }

double dot3(const double* a, const double* b)
{
    return std::fma(a[2], b[2], std::fma(a[0], b[0], a[1] * b[1]));
}

}  // namespace

namespace disco {

void release_prismatic_buffers(disco_ctx* ctx)
{
    if (ctx->d_slab_scratch) (void)hipFree(ctx->d_slab_scratch);
    if (ctx->h_slab_stage) (void)hipHostFree(ctx->h_slab_stage);
    ctx->d_slab_scratch = nullptr;
    ctx->h_slab_stage = nullptr;
    ctx->cap_slab_scratch = 0;
}

}  // namespace disco

extern "C" {

int disco_hip_matter_type(const disco_loop_state* loop)
{
    if (!loop) return -1;
    const double mn = dot3(loop->m, loop->n);
    if (std::fabs(mn) < 1e-6) return DISCO_MATTER_NONE;
    return mn > 0.0 ? DISCO_MATTER_VACANCY : DISCO_MATTER_INTERSTITIAL;
}

int disco_hip_prismatic_slab(disco_ctx* ctx, const disco_loop_state* loop, double spacing,
                             long long n, const double* d_x, const double* d_y,
                             const double* d_z, double lx, double ly, double lz, int* d_atoms,
                             int* d_side, long long cap, disco_slab_info* info, void* stream)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (!loop || !info) return fail(ctx, DISCO_ERR_ARG, "null loop or info");
    std::memset(info, 0, sizeof(*info));
    if (loop->shape < DISCO_SHAPE_CIRCLE || loop->shape > DISCO_SHAPE_PARALLELOGRAM)
        return fail(ctx, DISCO_ERR_ARG, "unknown loop shape");
    if (loop->shape == DISCO_SHAPE_POLYGON &&
        (loop->nvert < 3 || loop->nvert > DISCO_MAX_STATE_VERT))
        return fail(ctx, DISCO_ERR_ARG, "polygon vertex count out of range");
    if (!(spacing > 0.0) || !std::isfinite(spacing))
        return fail(ctx, DISCO_ERR_ARG, "plane spacing must be positive");
    if (!(lx > 0.0) || !(ly > 0.0) || !(lz > 0.0))
        return fail(ctx, DISCO_ERR_ARG, "box lengths must be positive");
    if (n < 0 || cap < 0) return fail(ctx, DISCO_ERR_ARG, "negative size");
    if (n > 0x7fffffffll) return fail(ctx, DISCO_ERR_TOO_LARGE, "atom indices are int");

    // matterType, then the slab height in planes (Code/prismatic.cpp:37-53)
    info->type = disco_hip_matter_type(loop);
    if (info->type == DISCO_MATTER_NONE) return DISCO_OK;
    for (int i = 0; i < 3; i++) info->b[i] = loop->bmag * loop->m[i];
    info->height = std::fabs(dot3(info->b, loop->n));
    info->spacing = spacing;
    const double ratio = info->height / spacing;
    info->planes = (int)std::lround(ratio);
    if (info->planes < 1 || std::fabs(ratio - info->planes) > 1e-6)
        return fail(ctx, DISCO_ERR_ARG,
                    "|b.n| is not a whole number of atomic-plane spacings: this b cannot close "
                    "on this habit plane");
    if (n == 0) return DISCO_OK;
    if (!d_x || !d_y || !d_z) return fail(ctx, DISCO_ERR_ARG, "null device array");
    if (cap > 0 && (!d_atoms || !d_side)) return fail(ctx, DISCO_ERR_ARG, "null output array");

    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    // pinned staging: the loop copy (H2D) and the slab total (D2H)
    if (!ctx->h_slab_stage)
        DISCO_HIP_TRY(ctx, hipHostMalloc(&ctx->h_slab_stage, sizeof(SlabLoop) + 256,
                                         hipHostMallocDefault));
    SlabLoop& P = *static_cast<SlabLoop*>(ctx->h_slab_stage);
    unsigned long long* h_total =
        reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->h_slab_stage) +
                                              ((sizeof(SlabLoop) + 63) & ~(size_t)63));
    std::memset(&P, 0, sizeof(P));
    P.s = *loop;
    P.cuv = dot3(loop->u, loop->v);
    P.det = std::fma(-P.cuv, P.cuv, 1.0);
    P.t_lo = -std::floor(info->planes / 2.);
    P.t_hi = std::ceil(info->planes / 2.);
    P.spacing = spacing;
    P.lx = lx;
    P.ly = ly;
    P.lz = lz;

    const long long nb = (n + kSlabTile - 1) / kSlabTile;
    // scratch: loop | total | tile counts | tile offsets | flags
    const size_t o_total = (sizeof(SlabLoop) + 255) & ~(size_t)255;
    const size_t o_counts = o_total + 256;
    const size_t o_offs = (o_counts + (size_t)nb * 4 + 255) & ~(size_t)255;
    const size_t o_flags = (o_offs + (size_t)nb * 8 + 255) & ~(size_t)255;
    const size_t need = o_flags + (size_t)nb * kSlabTile;
    int rc = ensure(ctx, &ctx->d_slab_scratch, &ctx->cap_slab_scratch, need);
    if (rc) return rc;
    char* sc = static_cast<char*>(ctx->d_slab_scratch);
    SlabLoop* d_loop = reinterpret_cast<SlabLoop*>(sc);
    unsigned long long* d_total = reinterpret_cast<unsigned long long*>(sc + o_total);
    unsigned* d_counts = reinterpret_cast<unsigned*>(sc + o_counts);
    unsigned long long* d_offs = reinterpret_cast<unsigned long long*>(sc + o_offs);
    signed char* d_flags = reinterpret_cast<signed char*>(sc + o_flags);
    hipStream_t s = (hipStream_t)stream;
    DISCO_HIP_TRY(ctx, hipMemcpyAsync(d_loop, &P, sizeof(P), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(slab_flag_kernel, dim3((unsigned)nb), dim3(kSlabBlock), 0, s, d_loop, n,
                       d_x, d_y, d_z, d_flags, d_counts);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    hipLaunchKernelGGL(slab_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, d_counts, d_offs, nb,
                       d_total);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    if (cap > 0) {
        hipLaunchKernelGGL(slab_scatter_kernel, dim3((unsigned)nb), dim3(kSlabBlock), 0, s, n,
                           d_flags, d_offs, d_atoms, d_side, cap);
        DISCO_HIP_TRY(ctx, hipGetLastError());
    }
    DISCO_HIP_TRY(ctx, hipMemcpyAsync(h_total, d_total, sizeof(*h_total), hipMemcpyDeviceToHost, s));
    DISCO_HIP_TRY(ctx, hipStreamSynchronize(s));
    info->count = (long long)*h_total;
    if (info->count > cap)
        return fail(ctx, DISCO_ERR_TOO_LARGE, "slab larger than the output capacity");
    return DISCO_OK;
}

int disco_hip_slab_filter(disco_ctx* ctx, long long natoms, const int* d_grain, int grain,
                          const int* d_twin, int region, int* d_atoms, int* d_side,
                          long long* count, void* stream)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (!count || *count < 0) return fail(ctx, DISCO_ERR_ARG, "null or negative count");
    if (natoms < 0) return fail(ctx, DISCO_ERR_ARG, "negative atom count");
    const long long n = *count;
    if (n == 0 || (!d_grain && !d_twin)) return DISCO_OK;
    if (!d_atoms || !d_side) return fail(ctx, DISCO_ERR_ARG, "null slab arrays");
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    unsigned long long* h_total;
    int rc = pinned_total(ctx, &h_total);
    if (rc) return rc;
    Compaction c;
    rc = compaction_scratch(ctx, n, (size_t)n * 8, &c);
    if (rc) return rc;
    int* tmp_atoms = reinterpret_cast<int*>(c.extra);
    int* tmp_side = tmp_atoms + n;
    hipStream_t s = (hipStream_t)stream;
    rc = launch_compaction(ctx, n, KeepRegion{d_atoms, d_grain, d_twin, grain, region, natoms},
                           MoveSlab{d_atoms, d_side, tmp_atoms, tmp_side}, c, s);
    if (rc) return rc;
    DISCO_HIP_TRY(ctx, hipMemcpyAsync(h_total, c.total, sizeof(*h_total), hipMemcpyDeviceToHost, s));
    DISCO_HIP_TRY(ctx, hipStreamSynchronize(s));
    const long long kept = (long long)*h_total;
    if (kept > 0) {
        DISCO_HIP_TRY(ctx, hipMemcpyAsync(d_atoms, tmp_atoms, (size_t)kept * 4,
                                          hipMemcpyDeviceToDevice, s));
        DISCO_HIP_TRY(ctx, hipMemcpyAsync(d_side, tmp_side, (size_t)kept * 4,
                                          hipMemcpyDeviceToDevice, s));
        DISCO_HIP_TRY(ctx, hipStreamSynchronize(s));
    }
    *count = kept;
    return DISCO_OK;
}

int disco_hip_slab_mark_removed(disco_ctx* ctx, long long natoms, long long count,
                                const int* d_atoms, unsigned char* d_removed, void* stream)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (count < 0 || natoms < 0) return fail(ctx, DISCO_ERR_ARG, "negative count");
    if (count == 0) return DISCO_OK;
    if (!d_atoms || !d_removed) return fail(ctx, DISCO_ERR_ARG, "null device array");
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(mark_removed_kernel, dim3(blocks256(count)), dim3(256), 0,
                       (hipStream_t)stream, natoms, count, d_atoms, d_removed);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    return DISCO_OK;
}

int disco_hip_slab_copies(disco_ctx* ctx, long long count, const int* d_atoms, const int* d_side,
                          const double shift[6], int* d_copy_atom, double* d_copy_shift,
                          void* stream)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (count < 0 || !shift) return fail(ctx, DISCO_ERR_ARG, "negative count or null shift");
    if (count == 0) return DISCO_OK;
    if (!d_atoms || !d_side || !d_copy_atom || !d_copy_shift)
        return fail(ctx, DISCO_ERR_ARG, "null device array");
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(slab_copies_kernel, dim3(blocks256(count)), dim3(256), 0,
                       (hipStream_t)stream, count, d_atoms, d_side, shift[0], shift[1], shift[2],
                       shift[3], shift[4], shift[5], d_copy_atom, d_copy_shift);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    return DISCO_OK;
}

int disco_hip_drop_atoms(disco_ctx* ctx, long long n, const unsigned char* d_removed,
                         const double* d_x, const double* d_y, const double* d_z,
                         const int* d_label, double* d_xo, double* d_yo, double* d_zo,
                         int* d_label_o, long long ncopies, const int* d_copy_atom,
                         const double* d_copy_shift, int* d_copy_atom_o, double* d_copy_shift_o,
                         long long* n_kept, long long* ncopies_kept, void* stream)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (!n_kept || !ncopies_kept) return fail(ctx, DISCO_ERR_ARG, "null output count");
    *n_kept = 0;
    *ncopies_kept = 0;
    if (n < 0 || ncopies < 0) return fail(ctx, DISCO_ERR_ARG, "negative size");
    if (n > 0x7fffffffll) return fail(ctx, DISCO_ERR_TOO_LARGE, "atom indices are int");
    if (n == 0) {
        if (ncopies > 0) return fail(ctx, DISCO_ERR_ARG, "copies of atoms that do not exist");
        return DISCO_OK;
    }
    if (!d_removed || !d_x || !d_y || !d_z || !d_xo || !d_yo || !d_zo ||
        (d_label && !d_label_o))
        return fail(ctx, DISCO_ERR_ARG, "null device array");
    if (ncopies > 0 && (!d_copy_atom || !d_copy_shift || !d_copy_atom_o || !d_copy_shift_o))
        return fail(ctx, DISCO_ERR_ARG, "null copy array");
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    unsigned long long* h_total;
    int rc = pinned_total(ctx, &h_total);
    if (rc) return rc;
    Compaction c;
    rc = compaction_scratch(ctx, std::max(n, ncopies), (size_t)n * 4, &c);
    if (rc) return rc;
    int* new_index = reinterpret_cast<int*>(c.extra);
    hipStream_t s = (hipStream_t)stream;
    rc = launch_compaction(ctx, n, KeepAtom{d_removed},
                           MoveAtom{d_x, d_y, d_z, d_label, d_xo, d_yo, d_zo, d_label_o, new_index},
                           c, s);
    if (rc) return rc;
    DISCO_HIP_TRY(ctx, hipMemcpyAsync(h_total, c.total, sizeof(*h_total), hipMemcpyDeviceToHost, s));
    DISCO_HIP_TRY(ctx, hipStreamSynchronize(s));
    *n_kept = (long long)*h_total;
    if (ncopies > 0) {
        rc = launch_compaction(ctx, ncopies, KeepCopy{d_copy_atom, new_index, n},
                               MoveCopy{d_copy_atom, new_index, d_copy_shift, d_copy_atom_o,
                                        d_copy_shift_o},
                               c, s);
        if (rc) return rc;
        DISCO_HIP_TRY(ctx, hipMemcpyAsync(h_total, c.total, sizeof(*h_total),
                                          hipMemcpyDeviceToHost, s));
        DISCO_HIP_TRY(ctx, hipStreamSynchronize(s));
        *ncopies_kept = (long long)*h_total;
    }
    return DISCO_OK;
}

int disco_hip_append_copies(disco_ctx* ctx, long long n, long long ncopies,
                            const int* d_copy_atom, const double* d_copy_shift, double* d_x,
                            double* d_y, double* d_z, double* d_ux, double* d_uy, double* d_uz,
                            int* d_label, void* stream)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (n < 0 || ncopies < 0) return fail(ctx, DISCO_ERR_ARG, "negative size");
    if (ncopies == 0) return DISCO_OK;
    if (n == 0) return fail(ctx, DISCO_ERR_ARG, "copies of atoms that do not exist");
    if (n + ncopies > 0x7fffffffll) return fail(ctx, DISCO_ERR_TOO_LARGE, "atom indices are int");
    if (!d_copy_atom || !d_copy_shift || !d_x || !d_y || !d_z || !d_ux || !d_uy || !d_uz)
        return fail(ctx, DISCO_ERR_ARG, "null device array");
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipLaunchKernelGGL(append_copies_kernel, dim3(blocks256(ncopies)), dim3(256), 0,
                       (hipStream_t)stream, n, ncopies, d_copy_atom, d_copy_shift, d_x, d_y, d_z,
                       d_ux, d_uy, d_uz, d_label);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    return DISCO_OK;
}

}  // extern "C"
