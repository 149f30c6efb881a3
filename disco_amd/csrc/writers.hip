// writers.hip -- the output stage on gfx950: final positions and the text of
// the XYZ / LAMMPS writers (include/disco_output.h).
//
// Reference behaviour replaced: the final-position loop and the two writers of
// Code/main.cpp (:797-845 and writeLAMMPS, :50-135).  Those run on one host
// thread and format every coordinate through an iostream; a 4M-atom sample is
// ~180 MB of text.  Here:
//   * final_positions_kernel: one atom per lane, 48 B read and 24 B written per
//     atom (HBM-bound elementwise pass);
//   * the text of atom lines in three launches: text_count_kernel (bytes per
//     256-atom block), text_scan_kernel (block offsets), text_write_kernel
//     (each lane formats its line into LDS at its block-local offset, then the
//     block stores its contiguous text with aligned dword stores: the global
//     writes are coalesced whatever the line lengths);
//   * a coordinate is printed exactly as glibc prints "%.8f": N = |v| 10^8
//     rounded half-to-even from the exact binary value (m 5^8 2^(e+8) in 128-bit
//     integer arithmetic), then N's decimal digits -- no floating-point
//     rounding step, so the text equals the host's for every double;
//   * files: the header is formatted on the host (iostreams, as the reference),
//     the lines are formatted per chunk of 2^20 atoms on the device and copied
//     through two pinned 16 MiB buffers, the copy of one piece overlapping the
//     fwrite of the previous one.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <sstream>
#include <string>

#include "ctx_internal.h"
#include "disco_output.h"

#pragma clang fp contract(off)

using namespace disco;

namespace {

constexpr int kTextBlock = 256;
constexpr int kMaxLine = 128;            // bytes: bound of any line (checked below)
constexpr long long kChunkAtoms = 1 << 20;
constexpr size_t kStageBytes = 16u << 20;
constexpr int kScanBlock = 1024;

struct TextArgs {
    const double* x;
    const double* y;
    const double* z;
    const int* grain;
    const int* twin;
    long long n;
    long long first_id;
    double f[9];
    char element[16];
    int elen;
};

// ------------------------------------------------------------- final positions
__global__ __launch_bounds__(256) void final_positions_kernel(
    long long n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ ux, const double* __restrict__ uy,
    const double* __restrict__ uz, double dx, double dy, double dz, double hx, double hy,
    double hz, double u00, double u01, double u02, double u10, double u11, double u12, double u20,
    double u21, double u22, double* __restrict__ fx, double* __restrict__ fy,
    double* __restrict__ fz)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double px = x[k] + ux[k];
    double py = y[k] + uy[k];
    double pz = z[k] + uz[k];
    if (px > hx) px -= dx; else if (px < -hx) px += dx;
    if (py > hy) py -= dy; else if (py < -hy) py += dy;
    if (pz > hz) pz -= dz; else if (pz < -hz) pz += dz;
    const double sx = px + hx, sy = py + hy, sz = pz + hz;
    // a*b + c*d + e*f as the reference's host build evaluates it (GCC fuses it
    // into fma(e, f, fma(a, b, c*d)))
    fx[k] = fma(u02, sz, fma(u00, sx, u01 * sy));
    fy[k] = fma(u12, sz, fma(u10, sx, u11 * sy));
    fz[k] = fma(u22, sz, fma(u20, sx, u21 * sy));
}

// ------------------------------------------------------------- fixed, 8 decimals
struct Fix8 {
    unsigned long long ip;   // integer part
    unsigned fp;             // the 8 decimals as an integer
    int neg;                 // leading '-' (sign bit set, -0.0 included)
    int kind;                // 0 finite, 1 inf, 2 nan, 3 out of range
};

__device__ __forceinline__ Fix8 fix8(double v)
{
    Fix8 r;
    const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    r.neg = (int)(bits >> 63);
    const int ex = (int)((bits >> 52) & 0x7ff);
    const unsigned long long frac = bits & 0xfffffffffffffull;
    r.ip = 0;
    r.fp = 0;
    if (ex == 0x7ff) {
        r.kind = frac ? 2 : 1;
        return r;
    }
    if (!(fabs(v) < 1e10)) {
        r.kind = 3;
        return r;
    }
    r.kind = 0;
    const unsigned long long m = ex ? (frac | 0x10000000000000ull) : frac;
    const int s = (ex ? ex - 1075 : -1074) + 8;        // |v| 10^8 = m 5^8 2^s
    const unsigned __int128 P = (unsigned __int128)m * 390625u;   // < 2^72
    unsigned long long N;
    if (s >= 0) {
        N = (unsigned long long)(P << s);              // exact, < 10^18
    } else if (s < -72) {
        N = 0;                                         // below one half
    } else {
        const int sh = -s;
        const unsigned __int128 q = P >> sh;
        const unsigned __int128 rem = P - (q << sh);
        const unsigned __int128 half = (unsigned __int128)1 << (sh - 1);
        N = (unsigned long long)q + ((rem > half || (rem == half && (q & 1))) ? 1u : 0u);
    }
    r.ip = N / 100000000ull;
    r.fp = (unsigned)(N - r.ip * 100000000ull);
    return r;
}

__device__ __forceinline__ int ndigits(unsigned long long v)
{
    int d = 1;
    unsigned long long p = 10;
#pragma unroll
    for (int i = 0; i < 19; i++) {
        d += v >= p ? 1 : 0;
        p = p * 10;
    }
    return d;
}

__device__ __forceinline__ int fix8_len(const Fix8& r)
{
    return r.kind ? 3 + r.neg : r.neg + ndigits(r.ip) + 9;
}

__device__ __forceinline__ int int_len(long long v)
{
    return (v < 0 ? 1 : 0) + ndigits(v < 0 ? (unsigned long long)(-(v + 1)) + 1ull
                                           : (unsigned long long)v);
}

__device__ __forceinline__ char* put_udec(char* p, unsigned long long v, int nd)
{
    if (v >> 32) {
        for (int i = nd - 1; i >= 0; i--) {
            p[i] = (char)('0' + (int)(v % 10));
            v /= 10;
        }
    } else {
        unsigned w = (unsigned)v;
        for (int i = nd - 1; i >= 0; i--) {
            p[i] = (char)('0' + (int)(w % 10u));
            w /= 10u;
        }
    }
    return p + nd;
}

__device__ __forceinline__ char* put_int(char* p, long long v)
{
    unsigned long long a = (unsigned long long)v;
    if (v < 0) {
        *p++ = '-';
        a = (unsigned long long)(-(v + 1)) + 1ull;
    }
    return put_udec(p, a, ndigits(a));
}

__device__ __forceinline__ char* put_fix8(char* p, const Fix8& r)
{
    if (r.neg) *p++ = '-';
    if (r.kind) {
        const char* w = r.kind == 2 ? "nan" : "inf";
        p[0] = w[0];
        p[1] = w[1];
        p[2] = w[2];
        return p + 3;
    }
    p = put_udec(p, r.ip, ndigits(r.ip));
    *p++ = '.';
    unsigned f = r.fp;
#pragma unroll
    for (int i = 7; i >= 0; i--) {
        p[i] = (char)('0' + (int)(f % 10u));
        f /= 10u;
    }
    return p + 8;
}

// One atom's line: its values, their length, and the writer.
template <int FMT>
struct Line {
    Fix8 c[3];
    long long i0, i1;   // XYZ: grain + 1, twin; LAMMPS: id
    int len;
    int bad;

    __device__ __forceinline__ void load(const TextArgs& a, long long k)
    {
        const double x = a.x[k], y = a.y[k], z = a.z[k];
        if (FMT == DISCO_TEXT_XYZ) {
            c[0] = fix8(x);
            c[1] = fix8(y);
            c[2] = fix8(z);
            len = a.elen + 4 + fix8_len(c[0]) + fix8_len(c[1]) + fix8_len(c[2]);
            i0 = 0;
            i1 = 0;
            if (a.grain) {
                i0 = (long long)a.grain[k] + 1;
                len += 1 + int_len(i0);
            }
            if (a.twin) {
                i1 = a.twin[k];
                len += 1 + int_len(i1);
            }
        } else {
            // e . r as the reference's host build evaluates it
            c[0] = fix8(fma(a.f[2], z, fma(a.f[0], x, a.f[1] * y)));
            c[1] = fix8(fma(a.f[5], z, fma(a.f[3], x, a.f[4] * y)));
            c[2] = fix8(fma(a.f[8], z, fma(a.f[6], x, a.f[7] * y)));
            i0 = a.first_id + k;
            i1 = 0;
            len = int_len(i0) + 6 + fix8_len(c[0]) + fix8_len(c[1]) + fix8_len(c[2]);
        }
        bad = (c[0].kind == 3) | (c[1].kind == 3) | (c[2].kind == 3);
    }

    __device__ __forceinline__ void write(const TextArgs& a, char* p) const
    {
        if (FMT == DISCO_TEXT_XYZ) {
            for (int i = 0; i < a.elen; i++) p[i] = a.element[i];
            p += a.elen;
            for (int j = 0; j < 3; j++) {
                *p++ = ' ';
                p = put_fix8(p, c[j]);
            }
            if (a.grain) {
                *p++ = ' ';
                p = put_int(p, i0);
            }
            if (a.twin) {
                *p++ = ' ';
                p = put_int(p, i1);
            }
        } else {
            p = put_int(p, i0);
            p[0] = ' ';
            p[1] = '1';
            p += 2;
            for (int j = 0; j < 3; j++) {
                *p++ = ' ';
                p = put_fix8(p, c[j]);
            }
        }
        *p = '\n';
    }
};

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

template <int FMT>
__global__ __launch_bounds__(kTextBlock) void text_count_kernel(TextArgs a,
                                                                unsigned* __restrict__ block_bytes,
                                                                unsigned* __restrict__ flags)
{
    __shared__ int wsum[kTextBlock / 64];
    const long long k = (long long)blockIdx.x * kTextBlock + threadIdx.x;
    int len = 0, bad = 0;
    if (k < a.n) {
        Line<FMT> L;
        L.load(a, k);
        len = L.len;
        bad = L.bad;
    }
    if (bad) atomicOr(flags, 1u);
    const int s = wave_sum(len);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kTextBlock / 64; w++) t += wsum[w];
        block_bytes[blockIdx.x] = (unsigned)t;
    }
}

// Exclusive scan of the block byte counts (one workgroup; nb any size).
__global__ __launch_bounds__(kScanBlock) void text_scan_kernel(
    const unsigned* __restrict__ block_bytes, unsigned long long* __restrict__ block_off, int nb,
    unsigned long long* __restrict__ total)
{
    __shared__ unsigned long long s[kScanBlock];
    unsigned long long carry = 0;
    const int t = threadIdx.x;
    for (int base = 0; base < nb; base += kScanBlock) {
        const int i = base + t;
        const unsigned long long v = i < nb ? block_bytes[i] : 0ull;
        s[t] = v;
        __syncthreads();
        for (int o = 1; o < kScanBlock; o <<= 1) {
            const unsigned long long w = t >= o ? s[t - o] : 0ull;
            __syncthreads();
            s[t] += w;
            __syncthreads();
        }
        if (i < nb) block_off[i] = carry + s[t] - v;
        carry += s[kScanBlock - 1];
        __syncthreads();
    }
    if (t == 0) *total = carry;
}

template <int FMT>
__global__ __launch_bounds__(kTextBlock) void text_write_kernel(
    TextArgs a, const unsigned long long* __restrict__ block_off, char* __restrict__ out)
{
    __shared__ __align__(16) char buf[kTextBlock * kMaxLine + 16];
    __shared__ int wsum[kTextBlock / 64];
    const long long k = (long long)blockIdx.x * kTextBlock + threadIdx.x;
    Line<FMT> L;
    L.len = 0;
    if (k < a.n) L.load(a, k);

    // block-exclusive offset of this lane's line
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = L.len;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o);
        if (lane >= o) inc += v;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kTextBlock / 64; w++) {
        base += w < wave ? wsum[w] : 0;
        total += wsum[w];
    }
    if (L.len) L.write(a, buf + base + inc - L.len);
    __syncthreads();

    // the block's text: head bytes to a 4-byte boundary, dwords, tail bytes
    char* dst = out + block_off[blockIdx.x];
    int head = (int)((4u - ((unsigned)(uintptr_t)dst & 3u)) & 3u);
    if (head > total) head = total;
    if ((int)threadIdx.x < head) dst[threadIdx.x] = buf[threadIdx.x];
    const int words = (total - head) >> 2;
    const unsigned* s32 = reinterpret_cast<const unsigned*>(buf);
    unsigned* d32 = reinterpret_cast<unsigned*>(dst + head);
    const unsigned shift = (unsigned)head & 3u;
    for (int w = threadIdx.x; w < words; w += kTextBlock) {
        const int src = (head >> 2) + w;   // head < 4: word index of byte head + 4w
        d32[w] = __builtin_amdgcn_alignbyte(s32[src + 1], s32[src], shift);
    }
    const int tail0 = head + 4 * words;
    const int tail = total - tail0;
    if ((int)threadIdx.x < tail) dst[tail0 + threadIdx.x] = buf[tail0 + threadIdx.x];
}

// ------------------------------------------------------------- host helpers
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

int fill_args(disco_ctx* ctx, const disco_text_format* fmt, long long n, const double* fx,
              const double* fy, const double* fz, TextArgs* a)
{
    if (!fmt) return fail(ctx, DISCO_ERR_ARG, "null text format");
    if (fmt->format != DISCO_TEXT_XYZ && fmt->format != DISCO_TEXT_LAMMPS)
        return fail(ctx, DISCO_ERR_ARG, "unknown text format");
    if (n < 0) return fail(ctx, DISCO_ERR_ARG, "negative size");
    if (n > 0 && (!fx || !fy || !fz)) return fail(ctx, DISCO_ERR_ARG, "null device array");
    std::memset(a, 0, sizeof(*a));
    a->x = fx;
    a->y = fy;
    a->z = fz;
    a->n = n;
    if (fmt->format == DISCO_TEXT_XYZ) {
        const size_t el = strnlen(fmt->element, sizeof(fmt->element));
        if (el == sizeof(fmt->element))
            return fail(ctx, DISCO_ERR_ARG, "element name longer than 15 characters");
        for (size_t i = 0; i < el; i++)
            if (fmt->element[i] == '\n')
                return fail(ctx, DISCO_ERR_ARG, "newline in the element name");
        std::memcpy(a->element, fmt->element, el);
        a->elen = (int)el;
        a->grain = fmt->d_grain;
        a->twin = fmt->d_twin;
    } else {
        if (fmt->first_id < 0 || fmt->first_id > (1ll << 62))
            return fail(ctx, DISCO_ERR_ARG, "first id out of range");
        a->first_id = fmt->first_id;
        for (int i = 0; i < 9; i++) {
            if (!std::isfinite(fmt->frame[i])) return fail(ctx, DISCO_ERR_ARG, "non-finite frame");
            a->f[i] = fmt->frame[i];
        }
    }
    return DISCO_OK;
}

// Counts, scans and (d_out != NULL) writes; synchronous on s.
int format_lines(disco_ctx* ctx, int format, const TextArgs& a, char* d_out, long long cap,
                 long long* nbytes, hipStream_t s)
{
    *nbytes = 0;
    if (a.n == 0) return DISCO_OK;
    const long long nb = (a.n + kTextBlock - 1) / kTextBlock;
    if (nb > 0x7fffffffll) return fail(ctx, DISCO_ERR_TOO_LARGE, "too many atoms in one call");
    // scratch: flags (u32) + total (u64) | block bytes (u32) | block offsets (u64)
    const size_t off_bytes = 16 + (size_t)nb * 4;
    const size_t need = ((off_bytes + 255) & ~(size_t)255) + (size_t)nb * 8;
    int rc = ensure(ctx, &ctx->d_text_scratch, &ctx->cap_text_scratch, need);
    if (rc) return rc;
    char* sc = static_cast<char*>(ctx->d_text_scratch);
    unsigned* flags = reinterpret_cast<unsigned*>(sc);
    unsigned long long* total = reinterpret_cast<unsigned long long*>(sc + 8);
    unsigned* block_bytes = reinterpret_cast<unsigned*>(sc + 16);
    unsigned long long* block_off =
        reinterpret_cast<unsigned long long*>(sc + ((off_bytes + 255) & ~(size_t)255));
    // (this path synchronizes below, so it is never inside a graph capture)
    DISCO_HIP_TRY(ctx, hipMemsetAsync(sc, 0, 16, s));
    const dim3 grid((unsigned)nb);
    const bool prof = ctx->text_profile != 0;
    if (prof) DISCO_HIP_TRY(ctx, hipEventRecord(ctx->text_ev[0], s));
    if (format == DISCO_TEXT_XYZ)
        hipLaunchKernelGGL(text_count_kernel<DISCO_TEXT_XYZ>, grid, dim3(kTextBlock), 0, s, a,
                           block_bytes, flags);
    else
        hipLaunchKernelGGL(text_count_kernel<DISCO_TEXT_LAMMPS>, grid, dim3(kTextBlock), 0, s, a,
                           block_bytes, flags);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    if (prof) DISCO_HIP_TRY(ctx, hipEventRecord(ctx->text_ev[1], s));
    hipLaunchKernelGGL(text_scan_kernel, dim3(1), dim3(kScanBlock), 0, s, block_bytes, block_off,
                       (int)nb, total);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    if (prof) DISCO_HIP_TRY(ctx, hipEventRecord(ctx->text_ev[2], s));
    unsigned long long host[2];
    DISCO_HIP_TRY(ctx, hipMemcpyAsync(host, sc, 16, hipMemcpyDeviceToHost, s));
    DISCO_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (host[0] & 1u)
        return fail(ctx, DISCO_ERR_TOO_LARGE, "a coordinate to print has |v| >= 1e10");
    *nbytes = (long long)host[1];
    if (prof) {
        DISCO_HIP_TRY(ctx, hipEventElapsedTime(&ctx->text_ms[0], ctx->text_ev[0], ctx->text_ev[1]));
        DISCO_HIP_TRY(ctx, hipEventElapsedTime(&ctx->text_ms[1], ctx->text_ev[1], ctx->text_ev[2]));
        ctx->text_ms[2] = 0.f;
    }
    if (!d_out) return DISCO_OK;
    if (*nbytes > cap) return fail(ctx, DISCO_ERR_TOO_LARGE, "text buffer too small");
    if (prof) DISCO_HIP_TRY(ctx, hipEventRecord(ctx->text_ev[2], s));
    if (format == DISCO_TEXT_XYZ)
        hipLaunchKernelGGL(text_write_kernel<DISCO_TEXT_XYZ>, grid, dim3(kTextBlock), 0, s, a,
                           block_off, d_out);
    else
        hipLaunchKernelGGL(text_write_kernel<DISCO_TEXT_LAMMPS>, grid, dim3(kTextBlock), 0, s, a,
                           block_off, d_out);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    if (prof) DISCO_HIP_TRY(ctx, hipEventRecord(ctx->text_ev[3], s));
    DISCO_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (prof)
        DISCO_HIP_TRY(ctx, hipEventElapsedTime(&ctx->text_ms[2], ctx->text_ev[2], ctx->text_ev[3]));
    return DISCO_OK;
}

// Gram-Schmidt of the box columns, with the products fused as the reference's
// host build fuses them (a*b + c*d + e*f -> fma(e, f, fma(a, b, c*d)),
// a - b*c -> fma(-b, c, a)).
double dot3(const double* a, const double* b) { return std::fma(a[2], b[2], std::fma(a[0], b[0], a[1] * b[1])); }

void lammps_frame(const double* Hp, double* box6, double* e)
{
    const double a1[3] = {Hp[0], Hp[3], Hp[6]};
    const double a2[3] = {Hp[1], Hp[4], Hp[7]};
    const double a3[3] = {Hp[2], Hp[5], Hp[8]};
    const double lx = std::sqrt(dot3(a1, a1));
    double* e1 = e;
    double* e2 = e + 3;
    double* e3 = e + 6;
    for (int i = 0; i < 3; i++) e1[i] = a1[i] / lx;
    const double xy = dot3(a2, e1);
    double r2[3];
    for (int i = 0; i < 3; i++) r2[i] = std::fma(-xy, e1[i], a2[i]);
    const double ly = std::sqrt(dot3(r2, r2));
    for (int i = 0; i < 3; i++) e2[i] = r2[i] / ly;
    const double xz = dot3(a3, e1);
    double r3[3];
    for (int i = 0; i < 3; i++) r3[i] = std::fma(-xz, e1[i], a3[i]);
    const double yz = dot3(r3, e2);
    double r3p[3];
    for (int i = 0; i < 3; i++) r3p[i] = std::fma(-yz, e2[i], r3[i]);
    const double lz = std::sqrt(dot3(r3p, r3p));
    for (int i = 0; i < 3; i++) e3[i] = r3p[i] / lz;
    box6[0] = lx;
    box6[1] = ly;
    box6[2] = lz;
    box6[3] = xy;
    box6[4] = xz;
    box6[5] = yz;
}

// Streams the lines of n atoms after `header` into path.
int write_file(disco_ctx* ctx, const char* path, const std::string& header,
               const disco_text_format& fmt, long long n, const double* fx, const double* fy,
               const double* fz, hipStream_t s)
{
    if (!path) return fail(ctx, DISCO_ERR_ARG, "null path");
    // validate the format once (chunks only move the pointers)
    TextArgs probe;
    int rc = fill_args(ctx, &fmt, n, fx, fy, fz, &probe);
    if (rc) return rc;
    const long long chunk = n < kChunkAtoms ? n : kChunkAtoms;
    rc = ensure(ctx, reinterpret_cast<void**>(&ctx->d_text), &ctx->cap_text,
                (size_t)(chunk > 0 ? chunk : 1) * kMaxLine);
    if (rc) return rc;
    for (int i = 0; i < 2; i++)
        if (!ctx->h_stage[i])
            DISCO_HIP_TRY(ctx, hipHostMalloc(reinterpret_cast<void**>(&ctx->h_stage[i]),
                                             kStageBytes, hipHostMallocDefault));
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(ctx, DISCO_ERR_IO, std::string("cannot open ") + path);
    auto io_fail = [&](const char* what) {
        std::fclose(f);
        return fail(ctx, DISCO_ERR_IO, std::string(what) + " " + path);
    };
    if (std::fwrite(header.data(), 1, header.size(), f) != header.size())
        return io_fail("cannot write");
    int piece = 0;
    size_t pending = 0;   // bytes of the previous piece, waiting in h_stage[(piece - 1) & 1]
    for (long long c0 = 0; c0 < n; c0 += chunk) {
        const long long cn = n - c0 < chunk ? n - c0 : chunk;
        disco_text_format cf = fmt;
        TextArgs a;
        if (cf.d_grain) cf.d_grain += c0;
        if (cf.d_twin) cf.d_twin += c0;
        cf.first_id = fmt.first_id + c0;
        rc = fill_args(ctx, &cf, cn, fx + c0, fy + c0, fz + c0, &a);
        if (rc) { std::fclose(f); return rc; }
        long long bytes = 0;
        rc = format_lines(ctx, cf.format, a, ctx->d_text, (long long)ctx->cap_text, &bytes, s);
        if (rc) { std::fclose(f); return rc; }
        for (long long o = 0; o < bytes; o += (long long)kStageBytes) {
            const size_t len = (size_t)(bytes - o < (long long)kStageBytes ? bytes - o
                                                                            : (long long)kStageBytes);
            char* h = ctx->h_stage[piece & 1];
            hipError_t e = hipMemcpyAsync(h, ctx->d_text + o, len, hipMemcpyDeviceToHost, s);
            if (e != hipSuccess) {
                std::fclose(f);
                return fail(ctx, DISCO_ERR_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
            }
            if (pending && std::fwrite(ctx->h_stage[(piece - 1) & 1], 1, pending, f) != pending)
                return io_fail("cannot write");
            e = hipStreamSynchronize(s);
            if (e != hipSuccess) {
                std::fclose(f);
                return fail(ctx, DISCO_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
            }
            pending = len;
            piece++;
        }
    }
    if (pending && std::fwrite(ctx->h_stage[(piece - 1) & 1], 1, pending, f) != pending)
        return io_fail("cannot write");
    if (std::fclose(f) != 0) return fail(ctx, DISCO_ERR_IO, std::string("cannot close ") + path);
    return DISCO_OK;
}

}  // namespace

namespace disco {

void release_output_buffers(disco_ctx* ctx)
{
    if (ctx->d_text_scratch) (void)hipFree(ctx->d_text_scratch);
    if (ctx->d_text) (void)hipFree(ctx->d_text);
    for (int i = 0; i < 2; i++)
        if (ctx->h_stage[i]) (void)hipHostFree(ctx->h_stage[i]);
    ctx->d_text_scratch = nullptr;
    ctx->d_text = nullptr;
    ctx->h_stage[0] = ctx->h_stage[1] = nullptr;
    ctx->cap_text_scratch = ctx->cap_text = 0;
    for (int i = 0; i < 4; i++)
        if (ctx->text_ev[i]) (void)hipEventDestroy(ctx->text_ev[i]);
    for (int i = 0; i < 4; i++) ctx->text_ev[i] = nullptr;
}

}  // namespace disco

extern "C" {

int disco_hip_final_positions(disco_ctx* ctx, long long n, const double* d_x, const double* d_y,
                              const double* d_z, const double* d_ux, const double* d_uy,
                              const double* d_uz, const double box[3], const double* U,
                              double* d_fx, double* d_fy, double* d_fz, void* stream)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (n < 0) return fail(ctx, DISCO_ERR_ARG, "negative size");
    if (!box) return fail(ctx, DISCO_ERR_ARG, "null box");
    for (int i = 0; i < 3; i++)
        if (!(box[i] > 0.0) || !std::isfinite(box[i]))
            return fail(ctx, DISCO_ERR_ARG, "box lengths must be positive and finite");
    if (n == 0) return DISCO_OK;
    if (!d_x || !d_y || !d_z || !d_ux || !d_uy || !d_uz || !d_fx || !d_fy || !d_fz)
        return fail(ctx, DISCO_ERR_ARG, "null device array");
    static const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const double* u = U ? U : I;
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const long long nb = (n + 255) / 256;
    if (nb > 0x7fffffffll) return fail(ctx, DISCO_ERR_TOO_LARGE, "too many atoms in one call");
    hipLaunchKernelGGL(final_positions_kernel, dim3((unsigned)nb), dim3(256), 0,
                       (hipStream_t)stream, n, d_x, d_y, d_z, d_ux, d_uy, d_uz, box[0], box[1],
                       box[2], box[0] / 2., box[1] / 2., box[2] / 2., u[0], u[1], u[2], u[3],
                       u[4], u[5], u[6], u[7], u[8], d_fx, d_fy, d_fz);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    return DISCO_OK;
}

int disco_hip_lammps_box(const double Hp[9], double box6[6], double frame9[9])
{
    if (!Hp || !box6 || !frame9) return DISCO_ERR_ARG;
    lammps_frame(Hp, box6, frame9);
    return DISCO_OK;
}

int disco_hip_format_text(disco_ctx* ctx, const disco_text_format* fmt, long long n,
                          const double* d_fx, const double* d_fy, const double* d_fz,
                          char* d_out, long long cap, long long* nbytes, void* stream)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (!nbytes) return fail(ctx, DISCO_ERR_ARG, "null nbytes");
    TextArgs a;
    int rc = fill_args(ctx, fmt, n, d_fx, d_fy, d_fz, &a);
    if (rc) return rc;
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    return format_lines(ctx, fmt->format, a, d_out, cap, nbytes, (hipStream_t)stream);
}

int disco_hip_text_profile(disco_ctx* ctx, int on)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (on && !ctx->text_ev[0]) {
        DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
        for (int i = 0; i < 4; i++) DISCO_HIP_TRY(ctx, hipEventCreate(&ctx->text_ev[i]));
    }
    ctx->text_profile = on ? 1 : 0;
    return DISCO_OK;
}

int disco_hip_text_kernel_ms(disco_ctx* ctx, double ms[3])
{
    if (!ctx || !ms) return DISCO_ERR_ARG;
    if (!ctx->text_profile) return fail(ctx, DISCO_ERR_STATE, "text profiling is off");
    for (int i = 0; i < 3; i++) ms[i] = ctx->text_ms[i];
    return DISCO_OK;
}

int disco_hip_write_xyz(disco_ctx* ctx, const char* path, long long n, const double* d_fx,
                        const double* d_fy, const double* d_fz, const double Hp[9],
                        const char* element, const int* d_grain, const int* d_twin, void* stream)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (!Hp || !element) return fail(ctx, DISCO_ERR_ARG, "null box or element");
    disco_text_format fmt;
    std::memset(&fmt, 0, sizeof(fmt));
    fmt.format = DISCO_TEXT_XYZ;
    if (std::strlen(element) >= sizeof(fmt.element))
        return fail(ctx, DISCO_ERR_ARG, "element name longer than 15 characters");
    std::strcpy(fmt.element, element);
    fmt.d_grain = d_grain;
    fmt.d_twin = d_twin;
    // header as the reference's stream prints it: count, then the Lattice
    // string (columns of Hp, default float format) and the property list
    std::ostringstream h;
    h << n << "\n" << "Lattice=\"";
    for (int c = 0; c < 3; c++) {
        if (c) h << "  ";
        h << Hp[c] << " " << Hp[3 + c] << " " << Hp[6 + c];
    }
    h << "\" Properties=species:S:1:pos:R:3" << (d_grain ? ":grain:I:1" : "")
      << (d_twin ? ":twin:I:1" : "") << "\n";
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    return write_file(ctx, path, h.str(), fmt, n, d_fx, d_fy, d_fz, (hipStream_t)stream);
}

int disco_hip_write_lammps(disco_ctx* ctx, const char* path, long long n, const double* d_fx,
                           const double* d_fy, const double* d_fz, const double Hp[9],
                           const char* element, double mass, void* stream)
{
    if (!ctx) return DISCO_ERR_ARG;
    if (!Hp || !element) return fail(ctx, DISCO_ERR_ARG, "null box or element");
    disco_text_format fmt;
    std::memset(&fmt, 0, sizeof(fmt));
    fmt.format = DISCO_TEXT_LAMMPS;
    fmt.first_id = 1;
    double b[6];
    lammps_frame(Hp, b, fmt.frame);
    std::ostringstream h;
    h << std::fixed << std::setprecision(8);
    h << "# LAMMPS data file generated by exec_cuda\n\n";
    h << n << " atoms\n\n";
    h << "1 atom types\n\n";
    h << "0.0 " << b[0] << " xlo xhi\n";
    h << "0.0 " << b[1] << " ylo yhi\n";
    h << "0.0 " << b[2] << " zlo zhi\n";
    if (std::fabs(b[3]) > 1e-8 || std::fabs(b[4]) > 1e-8 || std::fabs(b[5]) > 1e-8)
        h << b[3] << " " << b[4] << " " << b[5] << " xy xz yz\n";
    h << "\nMasses\n\n";
    h << "1 " << std::defaultfloat << std::setprecision(10) << mass << std::fixed
      << std::setprecision(8) << "  # " << element << "\n";
    h << "\nAtoms  # atomic\n\n";
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    return write_file(ctx, path, h.str(), fmt, n, d_fx, d_fy, d_fz, (hipStream_t)stream);
}

}  // extern "C"
