// grappa.hip -- in-plane GRAPPA on the GPU: the regularised least-squares fit on a reference scan
// (GrappaKernel::fit, src/kspace.rs:1540-1598, with solve_hermitian :1457-1491) and the synthesis of
// the missing phase-encode lines (GrappaKernel::apply, :1605-1633), in FP64 like the reference.
// Operation order follows the reference so the weights agree to round-off of the FP32 inputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <vector>

#include "internal.h"

#pragma clang fp contract(off)

namespace {

constexpr double kLambda = 1e-3;   // GRAPPA_LAMBDA (src/kspace.rs:1517)

struct Cd {
    double re, im;
};
__device__ __forceinline__ Cd cadd(Cd a, Cd b) { return Cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cd cmul(Cd a, Cd b) { return Cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cd cneg(Cd a) { return Cd{-a.re, -a.im}; }
__device__ __forceinline__ Cd crecip(Cd d) {
    const double m2 = d.re * d.re + d.im * d.im;
    return m2 < 1e-18 ? Cd{0.0, 0.0} : Cd{d.re / m2, -d.im / m2};
}

// Feature i at readout index kx from source lines s0, s1: (line, point, coil) = (i / 3nc, i / nc % 3,
// i % nc), readout point kx + {-1, 0, 1} wrapped (grappa_feature, src/kspace.rs:1521-1534).
__device__ __forceinline__ Cd feature(const float2* __restrict__ k, int nc, int nx, int ny, int s0, int s1, int kx,
                                      int i) {
    const int ci = i % nc, ki = (i / nc) % 3, li = i / (3 * nc);
    int v = (kx + ki - 1) % nx;
    if (v < 0) v += nx;
    const float2 z = k[((size_t)ci * ny + (li ? s1 : s0)) * nx + v];
    return Cd{(double)z.x, (double)z.y};
}

// One thread per entry of A^H A (nfeat^2) and A^H b (nc x nfeat) for offset d = blockIdx.y + 1,
// summed over the reference band in the reference's (b, kx) order.
__global__ void grappa_normal(const float2* __restrict__ k, int nc, int nx, int ny, int accel, int acs_lo, int acs_hi,
                              const uint8_t* __restrict__ lines, Cd* __restrict__ aha, Cd* __restrict__ ahb) {
    const int nfeat = 6 * nc;
    const int d = blockIdx.y + 1;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int nA = nfeat * nfeat;
    if (e >= nA + nc * nfeat) return;
    const bool is_a = e < nA;
    const int i = is_a ? e / nfeat : (e - nA) % nfeat;
    const int j = is_a ? e % nfeat : 0;
    const int co = is_a ? 0 : (e - nA) / nfeat;
    Cd acc{0.0, 0.0};
    for (int b = acs_lo; b < acs_hi; b++) {
        const int s0 = b, s1 = b + accel, tgt = b + d;
        if (s1 > acs_hi || tgt > acs_hi || !(lines[s0] && lines[s1] && lines[tgt])) continue;
        for (int kx = 0; kx < nx; kx++) {
            const Cd fi = feature(k, nc, nx, ny, s0, s1, kx, i);
            const Cd fic{fi.re, -fi.im};
            Cd other;
            if (is_a) {
                other = feature(k, nc, nx, ny, s0, s1, kx, j);
            } else {
                const float2 z = k[((size_t)co * ny + tgt) * nx + kx];
                other = Cd{(double)z.x, (double)z.y};
            }
            acc = cadd(acc, cmul(fic, other));
        }
    }
    const size_t base = (size_t)(d - 1);
    if (is_a)
        aha[base * nA + e] = acc;
    else
        ahb[base * nc * nfeat + (e - nA)] = acc;
}

// One block per offset: scale by the mean diagonal, add the Tikhonov term, then Gaussian elimination
// with partial pivoting on [A^H A | A^H b^T] and back substitution for every output coil
// (solve_hermitian, src/kspace.rs:1457-1491; the pivots depend on the matrix only, so one
// elimination serves every right-hand side exactly as separate solves do).
__global__ __launch_bounds__(256) void grappa_solve(Cd* __restrict__ aha, Cd* __restrict__ ahb, int nc,
                                                    Cd* __restrict__ weights) {
    __shared__ int s_piv;
    __shared__ double s_md;
    __shared__ Cd s_f[6 * MRS_MAX_COILS];
    const int n = 6 * nc, d = blockIdx.x;
    Cd* h = aha + (size_t)d * n * n;
    Cd* bb = ahb + (size_t)d * nc * n;   // [co][row]
    Cd* w = weights + (size_t)d * nc * n;
    if (threadIdx.x == 0) {
        double md = 0.0;
        for (int i = 0; i < n; i++) md += h[(size_t)i * n + i].re;
        s_md = md / (double)n;
    }
    __syncthreads();
    const double mean_diag = s_md;
    if (mean_diag <= 0.0) {
        for (int e = threadIdx.x; e < nc * n; e += blockDim.x) w[e] = Cd{0.0, 0.0};
        return;
    }
    const double sc = 1.0 / mean_diag;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        Cd v = h[e];
        v = Cd{v.re * sc, v.im * sc};
        if (e / n == e % n) v.re += kLambda;
        h[e] = v;
    }
    for (int e = threadIdx.x; e < nc * n; e += blockDim.x) bb[e] = Cd{bb[e].re * sc, bb[e].im * sc};
    __syncthreads();
    for (int col = 0; col < n; col++) {
        if (threadIdx.x == 0) {
            int piv = col;
            const Cd hc = h[(size_t)col * n + col];
            double best = sqrt(hc.re * hc.re + hc.im * hc.im);
            for (int r = col + 1; r < n; r++) {
                const Cd v = h[(size_t)r * n + col];
                const double m = sqrt(v.re * v.re + v.im * v.im);
                if (m > best) { best = m; piv = r; }
            }
            s_piv = piv;
        }
        __syncthreads();
        const int piv = s_piv;
        if (piv != col) {
            for (int c = threadIdx.x; c < n + nc; c += blockDim.x) {
                if (c < n) {
                    const Cd t = h[(size_t)col * n + c];
                    h[(size_t)col * n + c] = h[(size_t)piv * n + c];
                    h[(size_t)piv * n + c] = t;
                } else {
                    const int co = c - n;
                    const Cd t = bb[(size_t)co * n + col];
                    bb[(size_t)co * n + col] = bb[(size_t)co * n + piv];
                    bb[(size_t)co * n + piv] = t;
                }
            }
        }
        __syncthreads();
        const Cd dinv = crecip(h[(size_t)col * n + col]);
        for (int r = col + 1 + threadIdx.x; r < n; r += blockDim.x) s_f[r] = cmul(h[(size_t)r * n + col], dinv);
        __syncthreads();
        const int rows = n - col - 1, cols = n - col + nc;
        for (int e = threadIdx.x; e < rows * cols; e += blockDim.x) {
            const int r = col + 1 + e / cols, c = col + e % cols;
            const Cd f = s_f[r];
            if (c < n) {
                const Cd s = cmul(f, h[(size_t)col * n + c]);
                h[(size_t)r * n + c] = cadd(h[(size_t)r * n + c], cneg(s));
            } else {
                const int co = c - n;
                const Cd s = cmul(f, bb[(size_t)co * n + col]);
                bb[(size_t)co * n + r] = cadd(bb[(size_t)co * n + r], cneg(s));
            }
        }
        __syncthreads();
    }
    // back substitution, one thread per output coil, x kept in w[co][*]
    for (int co = threadIdx.x; co < nc; co += blockDim.x) {
        Cd* x = w + (size_t)co * n;
        for (int col = n - 1; col >= 0; col--) {
            Cd s = bb[(size_t)co * n + col];
            for (int c = col + 1; c < n; c++) s = cadd(s, cneg(cmul(h[(size_t)col * n + c], x[c])));
            x[col] = cmul(s, crecip(h[(size_t)col * n + col]));
        }
    }
}

// One thread per (unit, target line, kx, output coil): the synthesized sample, written in place.
struct ApplyRow {
    int ky, d;
};
__global__ void grappa_apply_k(float2* const* __restrict__ ks, int nunits, int nc, int nx, int ny, int accel,
                               const ApplyRow* __restrict__ rows, int nrows, const Cd* __restrict__ weights) {
    const int nfeat = 6 * nc;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per_unit = (size_t)nrows * nx * nc;
    if (e >= per_unit * nunits) return;
    const int u = (int)(e / per_unit);
    const size_t r = e % per_unit;
    const int co = (int)(r % nc), kx = (int)((r / nc) % nx), ri = (int)(r / ((size_t)nc * nx));
    const ApplyRow row = rows[ri];
    const int s0 = row.ky - row.d, s1 = row.ky - row.d + accel;
    float2* k = ks[u];
    const Cd* wd = weights + ((size_t)(row.d - 1) * nc + co) * nfeat;
    Cd v{0.0, 0.0};
    for (int i = 0; i < nfeat; i++) v = cadd(v, cmul(wd[i], feature(k, nc, nx, ny, s0, s1, kx, i)));
    k[((size_t)co * ny + row.ky) * nx + kx] = make_float2((float)v.re, (float)v.im);
}

int fail(char* err, size_t errlen, int code, const char* fmt, ...) {
    if (err && errlen) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(err, errlen, fmt, ap);
        va_end(ap);
    }
    return code;
}

}  // namespace

#define GTRY(expr)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            rc = fail(err, errlen, MRS_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));           \
            goto done;                                                                              \
        }                                                                                           \
    } while (0)

extern "C" int mrs_grappa_fit(const mrs_c32* ref_kspace, size_t n_coils, size_t nx, size_t ny, size_t accel,
                              size_t acs_lines, const uint8_t* lines, mrs_c64* weights, void* stream, char* err,
                              size_t errlen) {
    if (!ref_kspace || !lines || !weights) return fail(err, errlen, MRS_E_INVALID, "null argument");
    if (n_coils < 1 || n_coils > MRS_MAX_COILS) return fail(err, errlen, MRS_E_UNSUPPORTED, "1..%d coils", MRS_MAX_COILS);
    if (accel < 2) return MRS_OK;   // no offsets to fit
    if (nx < 1 || ny < 1) return fail(err, errlen, MRS_E_INVALID, "empty acquired matrix");
    const int nc = (int)n_coils, nfeat = 6 * nc, nd = (int)accel - 1;
    const long ys = (long)(ny / 2), half = (long)(acs_lines / 2);
    const int acs_lo = (int)std::max(ys - half, 0L), acs_hi = (int)std::min((size_t)(ys + half), ny - 1);
    hipStream_t st = (hipStream_t)stream;
    int rc = MRS_OK;
    Cd *aha = nullptr, *ahb = nullptr;
    uint8_t* dl = nullptr;
    const size_t na = (size_t)nd * nfeat * nfeat, nb = (size_t)nd * nc * nfeat;
    const int entries = nfeat * nfeat + nc * nfeat;
    GTRY(hipMalloc(&aha, na * sizeof(Cd)));
    GTRY(hipMalloc(&ahb, nb * sizeof(Cd)));
    GTRY(hipMalloc(&dl, ny));
    GTRY(hipMemcpyAsync(dl, lines, ny, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(grappa_normal, dim3((entries + 255) / 256, nd), dim3(256), 0, st,
                       reinterpret_cast<const float2*>(ref_kspace), nc, (int)nx, (int)ny, (int)accel, acs_lo, acs_hi,
                       dl, aha, ahb);
    GTRY(hipGetLastError());
    hipLaunchKernelGGL(grappa_solve, dim3(nd), dim3(256), 0, st, aha, ahb, nc, reinterpret_cast<Cd*>(weights));
    GTRY(hipGetLastError());
    GTRY(hipStreamSynchronize(st));
done:
    hipFree(aha);
    hipFree(ahb);
    hipFree(dl);
    return rc;
}

// ----------------------------------------------------------------------------- GRAPPA fit without allocation
// mrs_grappa_fit_noalloc keeps the fit's scratch (A^H A and A^H b per offset, then the reference scan's
// line flags) in one file-scope device array per GPU, `scratch`.  Host code never uses the array's
// name as an address: each device's copy is located once, on the first call with that device current,
// by hipGetSymbolAddress, and that pointer is what the kernels receive.  Calls that overlap on one GPU
// share its array and race; every value the kernels read is written earlier in the same call, so
// nothing needs clearing between calls.
static constexpr size_t kGrappaScratchBytes = 8u << 20;
static constexpr int kGrappaMaxLines = 1024;
static __device__ __align__(256) unsigned char scratch[kGrappaScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

struct LineFlags {
    uint32_t w[kGrappaMaxLines / 32];
};

// The reference scan's line flags into the scratch (a kernel argument, so no host copy).
__global__ void put_lines(uint8_t* __restrict__ dst, LineFlags f, int ny) {
    for (int i = threadIdx.x; i < ny; i += blockDim.x) dst[i] = (uint8_t)((f.w[i >> 5] >> (i & 31)) & 1u);
}

}  // namespace

extern "C" int mrs_grappa_fit_noalloc(const mrs_c32* ref_kspace, size_t n_coils, size_t nx, size_t ny, size_t accel,
                                      size_t acs_lines, const uint8_t* lines, mrs_c64* weights, void* stream,
                                      char* err, size_t errlen) {
    if (!ref_kspace || !lines || !weights) return fail(err, errlen, MRS_E_INVALID, "null argument");
    if (n_coils < 1 || n_coils > MRS_MAX_COILS) return fail(err, errlen, MRS_E_UNSUPPORTED, "1..%d coils", MRS_MAX_COILS);
    if (accel < 2) return MRS_OK;   // no offsets to fit
    if (nx < 1 || ny < 1) return fail(err, errlen, MRS_E_INVALID, "empty acquired matrix");
    const int nc = (int)n_coils, nfeat = 6 * nc, nd = (int)accel - 1;
    const size_t na = (size_t)nd * nfeat * nfeat, nb = (size_t)nd * nc * nfeat;
    const size_t need = (na + nb) * sizeof(Cd) + ny;
    if (ny > (size_t)kGrappaMaxLines || need > kGrappaScratchBytes)
        return fail(err, errlen, MRS_E_UNSUPPORTED,
                    "GRAPPA fit scratch of %zu bytes (%d coils, acceleration %zu, %zu lines) exceeds the no-alloc "
                    "limit of %zu bytes and %d lines", need, nc, accel, ny, kGrappaScratchBytes, kGrappaMaxLines);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(err, errlen, MRS_E_DEVICE, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(err, errlen, MRS_E_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            e = hipGetSymbolAddress(&g_scratch_ptr[dev], HIP_SYMBOL(scratch));
            if (e != hipSuccess) {
                g_scratch_ptr[dev] = nullptr;
                return fail(err, errlen, MRS_E_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            }
        }
        base = g_scratch_ptr[dev];
    }
    Cd* aha = reinterpret_cast<Cd*>(base);
    Cd* ahb = aha + na;
    uint8_t* dl = reinterpret_cast<uint8_t*>(ahb + nb);
    LineFlags f{};
    for (size_t i = 0; i < ny; i++)
        if (lines[i]) f.w[i >> 5] |= 1u << (i & 31);
    const long ys = (long)(ny / 2), half = (long)(acs_lines / 2);
    const int acs_lo = (int)std::max(ys - half, 0L), acs_hi = (int)std::min((size_t)(ys + half), ny - 1);
    hipStream_t st = (hipStream_t)stream;
    const int entries = nfeat * nfeat + nc * nfeat;
    hipLaunchKernelGGL(put_lines, dim3(1), dim3(256), 0, st, dl, f, (int)ny);
    hipLaunchKernelGGL(grappa_normal, dim3((entries + 255) / 256, nd), dim3(256), 0, st,
                       reinterpret_cast<const float2*>(ref_kspace), nc, (int)nx, (int)ny, (int)accel, acs_lo, acs_hi,
                       dl, aha, ahb);
    hipLaunchKernelGGL(grappa_solve, dim3(nd), dim3(256), 0, st, aha, ahb, nc, reinterpret_cast<Cd*>(weights));
    if ((e = hipGetLastError()) != hipSuccess) return fail(err, errlen, MRS_E_DEVICE, "launch: %s", hipGetErrorString(e));
    if ((e = hipStreamSynchronize(st)) != hipSuccess)
        return fail(err, errlen, MRS_E_DEVICE, "hipStreamSynchronize: %s", hipGetErrorString(e));
    return MRS_OK;
}

extern "C" int mrs_grappa_apply(mrs_c32* const* kspace, size_t n, size_t n_coils, size_t nx, size_t ny, size_t accel,
                                const uint8_t* acquired, const mrs_c64* weights, void* stream, char* err,
                                size_t errlen) {
    if (n == 0 || accel < 2) return MRS_OK;
    if (!kspace || !acquired || !weights) return fail(err, errlen, MRS_E_INVALID, "null argument");
    if (n_coils < 1 || n_coils > MRS_MAX_COILS) return fail(err, errlen, MRS_E_UNSUPPORTED, "1..%d coils", MRS_MAX_COILS);
    // the target lines (src/kspace.rs:1616-1623): offset d from the centre's phase, both sources acquired
    std::vector<ApplyRow> rows;
    const long ys = (long)(ny / 2);
    for (size_t d = 1; d < accel; d++)
        for (size_t ky = 0; ky < ny; ky++) {
            long m = ((long)ky - ys) % (long)accel;
            if (m < 0) m += (long)accel;
            if (m != (long)d || ky < d || ky + accel - d >= ny) continue;
            const size_t s0 = ky - d, s1 = ky - d + accel;
            if (!(acquired[s0] && acquired[s1])) continue;
            rows.push_back(ApplyRow{(int)ky, (int)d});
        }
    if (rows.empty()) return MRS_OK;
    hipStream_t st = (hipStream_t)stream;
    int rc = MRS_OK;
    ApplyRow* drows = nullptr;
    float2** dks = nullptr;
    const size_t total = n * rows.size() * nx * n_coils;
    GTRY(hipMalloc(&drows, rows.size() * sizeof(ApplyRow)));
    GTRY(hipMalloc(&dks, n * sizeof(float2*)));
    GTRY(hipMemcpyAsync(drows, rows.data(), rows.size() * sizeof(ApplyRow), hipMemcpyHostToDevice, st));
    GTRY(hipMemcpyAsync(dks, kspace, n * sizeof(float2*), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(grappa_apply_k, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, dks, (int)n,
                       (int)n_coils, (int)nx, (int)ny, (int)accel, drows, (int)rows.size(),
                       reinterpret_cast<const Cd*>(weights));
    GTRY(hipGetLastError());
    GTRY(hipStreamSynchronize(st));
done:
    hipFree(drows);
    hipFree(dks);
    return rc;
}
