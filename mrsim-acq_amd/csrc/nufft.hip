// nufft.hip -- the type-1 NUFFT y-stage of the reference's `kspace` feature build, in FP64 on the
// GPU: the comparator of the exact forward for uniform relaxation.
//
// mrs_nufft1: batched type-1 NUFFT (src/nufft.rs:25-113), F[k] = sum_j w_j exp(+i 2 pi k v_j / n), by
// gridding with the exponential-of-semicircle kernel psi(z) = exp(beta (sqrt(1 - z^2) - 1)) at
// oversampling 2 (fine grid nf = 2n, width w = clamp(ceil(-log10 tol) + 1, 2, 16), beta = 2.30 w):
// spread every source onto w fine cells, one inverse FFT of the fine grid, divide by the kernel's own
// transform at the n_out output frequencies.  One block per column; the fine grids of the column's
// weight vectors sit in LDS; the spreading accumulates with LDS f64 atomics (so the sums' order, and
// their last bits, vary from run to run); the FFT is a radix-2 decimation in LDS when nf is a power
// of two, else the needed bins are evaluated directly.
//
// mrs_build_coil_kspaces_nufft: build_coil_kspaces as the `kspace` feature builds it where its NUFFT
// y-stage applies (src/kspace.rs:778-831, 930-935): with t affine in the line index per parity and
// the readout polarity the parity's, the per-line y-sum of column x is one type-1 NUFFT of the
// sources displaced to v = (y - c) + sny tau fmap, per (compartment, parity) weight vector; each
// acquired line then combines the compartments with its uniform relaxation weights and takes the
// twiddle x-stage; spikes and noise as the reference adds them.  Everything after the host tables
// (timing, mask, per-line relaxation, twiddles, deapodisation) runs on the GPU in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "internal.h"

namespace mrs {

static constexpr double kTauN = 6.283185307179586476925286766559;

struct NufftShape {
    int nsrc, nw, nf, w, n_out;
    long k_lo;
    double beta, half;
    int pos_cols;   // columns share positions pos[col % pos_cols]
};

__device__ __forceinline__ double es_psi_d(double z, double beta) {
    return fabs(z) < 1.0 ? exp(beta * (sqrt(1.0 - z * z) - 1.0)) : 0.0;
}

// One block per column: spread, inverse FFT (or direct bins), deapodise.
__global__ __launch_bounds__(256) void nufft1_cols(const double* __restrict__ pos, const double2* __restrict__ wts,
                                                   const double* __restrict__ corr, double2* __restrict__ out,
                                                   NufftShape sh) {
    extern __shared__ double2 grid[];   // [nw][nf]
    const int col = blockIdx.x;
    const int nf = sh.nf, nw = sh.nw, nsrc = sh.nsrc, w = sh.w;
    for (int i = threadIdx.x; i < nw * nf; i += blockDim.x) grid[i] = make_double2(0.0, 0.0);
    __syncthreads();
    const double* p = pos + (size_t)(col % sh.pos_cols) * nsrc;
    const double2* wc = wts + (size_t)col * nw * nsrc;
    // spreading: one (source, tap) pair per thread iteration
    for (int t = threadIdx.x; t < nsrc * w; t += blockDim.x) {
        const int j = t / w, i = t % w;
        double pf = fmod(2.0 * p[j], (double)nf);
        if (pf < 0.0) pf += (double)nf;
        const long m0 = (long)floor(pf - sh.half) + 1;
        const double kv = es_psi_d(((double)m0 + (double)i - pf) / sh.half, sh.beta);
        long m = (m0 + i) % nf;
        if (m < 0) m += nf;
        for (int q = 0; q < nw; q++) {
            const double2 wv = wc[(size_t)q * nsrc + j];
            atomicAdd(&grid[q * nf + m].x, wv.x * kv);
            atomicAdd(&grid[q * nf + m].y, wv.y * kv);
        }
    }
    __syncthreads();
    const bool pow2 = (nf & (nf - 1)) == 0;
    if (pow2) {
        int lg = 0;
        while ((1 << lg) < nf) lg++;
        // bit-reversal permutation
        for (int t = threadIdx.x; t < nw * nf; t += blockDim.x) {
            const int q = t / nf, i = t % nf;
            const int r = (int)(__brev((unsigned)i) >> (32 - lg));
            if (i < r) {
                const double2 a = grid[q * nf + i];
                grid[q * nf + i] = grid[q * nf + r];
                grid[q * nf + r] = a;
            }
        }
        __syncthreads();
        // radix-2 stages of the unnormalised inverse transform (positive exponent)
        for (int half = 1; half < nf; half <<= 1) {
            for (int t = threadIdx.x; t < nw * (nf / 2); t += blockDim.x) {
                const int q = t / (nf / 2), k = t % (nf / 2);
                const int g = k / half, r = k % half;
                const int i = g * 2 * half + r, jx = i + half;
                double s, c;
                sincospi((double)r / (double)half, &s, &c);
                const double2 a = grid[q * nf + i], b = grid[q * nf + jx];
                const double tr = b.x * c - b.y * s, ti = b.x * s + b.y * c;
                grid[q * nf + i] = make_double2(a.x + tr, a.y + ti);
                grid[q * nf + jx] = make_double2(a.x - tr, a.y - ti);
            }
            __syncthreads();
        }
        for (int t = threadIdx.x; t < nw * sh.n_out; t += blockDim.x) {
            const int q = t / sh.n_out, jo = t % sh.n_out;
            long bin = (sh.k_lo + jo) % nf;
            if (bin < 0) bin += nf;
            const double2 v = grid[q * nf + bin];
            out[((size_t)col * nw + q) * sh.n_out + jo] = make_double2(v.x * corr[jo], v.y * corr[jo]);
        }
    } else {
        for (int t = threadIdx.x; t < nw * sh.n_out; t += blockDim.x) {
            const int q = t / sh.n_out, jo = t % sh.n_out;
            long bin = (sh.k_lo + jo) % nf;
            if (bin < 0) bin += nf;
            double ar = 0.0, ai = 0.0;
            for (int m = 0; m < nf; m++) {
                double s, c;
                sincospi(2.0 * (double)((bin * (long)m) % nf) / (double)nf, &s, &c);
                const double2 g = grid[q * nf + m];
                ar += g.x * c - g.y * s;
                ai += g.x * s + g.y * c;
            }
            out[((size_t)col * nw + q) * sh.n_out + jo] = make_double2(ar * corr[jo], ai * corr[jo]);
        }
    }
}

// The static set-up (src/kspace.rs:678-725) and the NUFFT sources of one slice: per voxel the
// displaced position v = (y - sys - yoff) + sny tau rate, and per coil the weights
// amp * comp_c * e^{i(phi0 + 2 pi rate t_centre)} (even lines) and the same times e^{i 2 pi rate delta}
// (odd lines), layout [coil][x][2c + parity][y].
struct NufftSlice {
    int snx, sny, nx, ny, Q, C, ox, oy, sxs, sys;
    double xoff, yoff, zc_c, tau_s, delta_s, t_centre, signal_scale, t_echo_s, eddy_phase;
    double grad[3];
    int do_dist, gradient_echo, do_eddy_phase;
};

__device__ double ring_sens_d(int coil, int n_coils, double x, double y, int nx, int ny) {
    if (n_coils <= 1) return 1.0;
    const double cx = (double)nx / 2.0, cy = (double)ny / 2.0;
    const double m = (double)max(nx, ny);
    const double r = 0.6 * m;
    const double ang = kTauN * (double)coil / (double)n_coils;
    const double px = cx + r * cos(ang), py = cy + r * sin(ang);
    const double dx = x - px, dy = y - py;
    const double sigma = 0.9 * m;
    return exp(-(dx * dx + dy * dy) / (2.0 * sigma * sigma)) + 0.15;
}

__global__ __launch_bounds__(256) void nufft_sources(const float* __restrict__ comps, const float* __restrict__ fmap,
                                                     const double* __restrict__ phase0, NufftSlice s,
                                                     double* __restrict__ pos, double2* __restrict__ wts) {
    const int P = s.snx * s.sny;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P) return;
    const int x = t % s.snx, y = t / s.snx;   // voxel i = x + snx*y
    const double f = (double)fmap[t];
    const double rate = s.do_dist ? f : 0.0;
    double p0 = phase0 ? phase0[t] : 0.0;
    if (s.do_eddy_phase) {
        const double xc = ((double)x - (double)s.sxs - s.xoff) / (double)s.ox;
        const double yc = ((double)y - (double)s.sys - s.yoff) / (double)s.oy;
        p0 += s.eddy_phase * (s.grad[0] * xc + s.grad[1] * yc + s.grad[2] * s.zc_c);
    }
    if (s.gradient_echo && s.do_dist) p0 += kTauN * f * s.t_echo_s;
    pos[(size_t)x * s.sny + y] = ((double)y - (double)s.sys - s.yoff) + (double)s.sny * s.tau_s * rate;
    const double ph = p0 + kTauN * rate * s.t_centre;
    const double c0 = cos(ph), s0 = sin(ph);
    const double phd = kTauN * rate * s.delta_s;
    const double cd = cos(phd), sd = sin(phd);
    const double c1 = c0 * cd - s0 * sd, s1 = c0 * sd + s0 * cd;
    const double xa = ((double)x - s.xoff) / (double)s.ox, ya = ((double)y - s.yoff) / (double)s.oy;
    const int nw = 2 * s.Q;
    for (int coil = 0; coil < s.C; coil++) {
        const double amp = s.signal_scale * ring_sens_d(coil, s.C, xa, ya, s.nx, s.ny);
        double2* wc = wts + ((size_t)coil * s.snx + x) * nw * s.sny;
        for (int c = 0; c < s.Q; c++) {
            const double a = amp * (double)comps[(size_t)c * P + t];
            wc[(size_t)(2 * c) * s.sny + y] = make_double2(a * c0, a * s0);
            wc[(size_t)(2 * c + 1) * s.sny + y] = make_double2(a * c1, a * s1);
        }
    }
}

// Per (acquired line, coil): g[x] = sum_c rel_c(line) row_c(x, line), then the twiddle x-stage and
// the 1/(snx sny) normalisation.  rows: [coil][x][2c + parity][ny]; tw: [table][kx][x].
__global__ __launch_bounds__(256) void nufft_lines(const double2* __restrict__ rows, const double* __restrict__ rel,
                                                   const double2* __restrict__ tw, const int8_t* __restrict__ table,
                                                   const uint8_t* __restrict__ acquired, NufftSlice s, double n_inv,
                                                   double2* __restrict__ k) {
    extern __shared__ double2 gl[];   // [snx]
    const int kyi = blockIdx.x, coil = blockIdx.y;
    double2* krow = k + ((size_t)coil * s.ny + kyi) * s.nx;
    if (!acquired[kyi]) {
        for (int kx = threadIdx.x; kx < s.nx; kx += blockDim.x) krow[kx] = make_double2(0.0, 0.0);
        return;
    }
    const int nw = 2 * s.Q, par = kyi % 2;
    for (int x = threadIdx.x; x < s.snx; x += blockDim.x) {
        double gr = 0.0, gi = 0.0;
        for (int c = 0; c < s.Q; c++) {
            const double2 v = rows[(((size_t)coil * s.snx + x) * nw + 2 * c + par) * s.ny + kyi];
            gr += rel[(size_t)kyi * s.Q + c] * v.x;
            gi += rel[(size_t)kyi * s.Q + c] * v.y;
        }
        gl[x] = make_double2(gr, gi);
    }
    __syncthreads();
    const double2* t = tw + (size_t)table[kyi] * s.nx * s.snx;
    for (int kx = threadIdx.x; kx < s.nx; kx += blockDim.x) {
        double ar = 0.0, ai = 0.0;
        for (int x = 0; x < s.snx; x++) {
            const double2 w = t[(size_t)kx * s.snx + x], g = gl[x];
            ar += g.x * w.x - g.y * w.y;
            ai += g.x * w.y + g.y * w.x;
        }
        krow[kx] = make_double2(ar * n_inv, ai * n_inv);
    }
}

// SplitMix64 (src/kspace.rs:60-77), counter-based: draw n of state s is mix(s + n*gamma).
__device__ __forceinline__ uint64_t sm_at(uint64_t s, uint64_t n) {
    uint64_t z = s + n * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double sm_unit(uint64_t v) { return ((double)(v >> 11) + 0.5) / 9007199254740992.0; }

// Spikes then thermal noise of one coil's k-space in FP64 (src/kspace.rs:997-1039): the spikes
// overwrite acquired samples picked by Rng(slice_seed ^ 0xA5A5_1234_5678_9ABC) with the coil's peak
// (first largest magnitude) times the amplitude; the noise adds to every acquired sample in index
// order from Rng((slice_seed ^ coil * 0x9E37_79B9) | 1), sample j using draws 4j+1..4j+4.
__global__ __launch_bounds__(256) void nufft_noise(double2* __restrict__ k, const int32_t* __restrict__ rows,
                                                   int nrows, int nx, int ny, uint64_t seed, int n_spikes,
                                                   double amp, double sigma) {
    __shared__ double bm[256];
    __shared__ int bi[256];
    const int coil = blockIdx.x;
    double2* K = k + (size_t)coil * nx * ny;
    const int nacq = nrows * nx;
    if (n_spikes > 0 && nacq > 0) {
        double best = -1.0;
        int bidx = 0x7fffffff;
        for (int i = threadIdx.x; i < nx * ny; i += blockDim.x) {
            const double m = sqrt(K[i].x * K[i].x + K[i].y * K[i].y);
            if (m > best || (m == best && i < bidx)) { best = m; bidx = i; }
        }
        bm[threadIdx.x] = best;
        bi[threadIdx.x] = bidx;
        __syncthreads();
        for (int off = blockDim.x / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) {
                const double m2 = bm[threadIdx.x + off];
                const int i2 = bi[threadIdx.x + off];
                if (m2 > bm[threadIdx.x] || (m2 == bm[threadIdx.x] && i2 < bi[threadIdx.x])) {
                    bm[threadIdx.x] = m2;
                    bi[threadIdx.x] = i2;
                }
            }
            __syncthreads();
        }
        // the peak is zero when every sample is (the reference's scan starts from a zero peak)
        const double2 peak = bm[0] > 0.0 ? K[bi[0]] : make_double2(0.0, 0.0);
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint64_t s = seed ^ 0xA5A5123456789ABCULL;
            for (int sp = 0; sp < n_spikes; sp++) {
                const int j = (int)(sm_at(s, (uint64_t)sp + 1) % (uint64_t)nacq);
                K[(size_t)rows[j / nx] * nx + (j % nx)] = make_double2(peak.x * amp, peak.y * amp);
            }
        }
        __syncthreads();
    }
    if (sigma > 0.0) {
        const uint64_t s0 = (seed ^ ((uint64_t)coil * 0x9E3779B9ULL)) | 1ULL;
        for (int j = threadIdx.x; j < nacq; j += blockDim.x) {
            const uint64_t b = 4ull * (uint64_t)j;
            const double u1 = sm_unit(sm_at(s0, b + 1)), u2 = sm_unit(sm_at(s0, b + 2));
            const double u3 = sm_unit(sm_at(s0, b + 3)), u4 = sm_unit(sm_at(s0, b + 4));
            double2& v = K[(size_t)rows[j / nx] * nx + (j % nx)];
            v.x += sqrt(-2.0 * log(u1)) * cos(kTauN * u2) * sigma;
            v.y += sqrt(-2.0 * log(u3)) * cos(kTauN * u4) * sigma;
        }
    }
}

// Width and shape of Nufft1::new (src/nufft.rs:43-66); the deapodisation on the host in FP64.
static int nufft_setup(size_t n, long k_lo, size_t n_out, double tol, NufftShape& sh, std::vector<double>& corr) {
    if (n_out == 0 || k_lo < -(long)(n / 2) || k_lo + (long)n_out - 1 > (long)(n / 2)) return -1;
    long wl = (long)std::ceil(-std::log10(tol)) + 1;
    const int w = (int)std::min<long>(16, std::max<long>(2, wl));
    sh.w = w;
    sh.beta = 2.30 * (double)w;
    sh.nf = (int)(2 * n);
    sh.half = (double)w / 2.0;
    sh.k_lo = k_lo;
    sh.n_out = (int)n_out;
    corr.resize(n_out);
    auto psi = [&](double z) { return std::fabs(z) < 1.0 ? std::exp(sh.beta * (std::sqrt(1.0 - z * z) - 1.0)) : 0.0; };
    for (size_t j = 0; j < n_out; j++) {
        const double k = (double)(k_lo + (long)j);
        double s = psi(0.0);
        for (double m = 1.0; m < sh.half; m += 1.0) s += 2.0 * psi(m / sh.half) * std::cos(kTauN * k * m / (double)sh.nf);
        corr[j] = 1.0 / s;
    }
    return 0;
}

static hipError_t launch_nufft1(const double* pos, const double2* w, const double* corr, double2* out, int ncols,
                                const NufftShape& sh, hipStream_t st) {
    const size_t lds = (size_t)sh.nw * sh.nf * sizeof(double2);
    hipError_t e = hipFuncSetAttribute((const void*)nufft1_cols, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(nufft1_cols, dim3(ncols), dim3(256), lds, st, pos, w, corr, out, sh);
    return hipGetLastError();
}

}  // namespace mrs

using namespace mrs;

static int nfail(char* err, size_t errlen, int code, const char* fmt, ...) {
    if (err && errlen) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(err, errlen, fmt, ap);
        va_end(ap);
    }
    return code;
}

#define NTRY(expr)                                                                                       \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) { rc = nfail(err, errlen, MRS_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); goto done; } \
    } while (0)

static constexpr size_t kMaxNufftLds = 160 * 1024;

extern "C" int mrs_nufft1(const double* pos, size_t pos_cols, const mrs_c64* weights, size_t nsrc, size_t nw,
                          size_t ncols, size_t n, long k_lo, size_t n_out, double tol, mrs_c64* out, void* stream,
                          char* err, size_t errlen) {
    if (!pos || !weights || !out) return nfail(err, errlen, MRS_E_INVALID, "null argument");
    if (ncols == 0 || nw == 0) return MRS_OK;
    if (pos_cols == 0 || ncols % pos_cols) return nfail(err, errlen, MRS_E_INVALID, "columns must share positions evenly");
    if (!(tol > 0.0)) return nfail(err, errlen, MRS_E_INVALID, "tolerance must be positive");
    NufftShape sh{};
    std::vector<double> corr;
    if (nufft_setup(n, k_lo, n_out, tol, sh, corr))
        return nfail(err, errlen, MRS_E_INVALID, "output band must sit inside the coarse Nyquist band");
    if (nw * (size_t)sh.nf * sizeof(double2) > kMaxNufftLds)
        return nfail(err, errlen, MRS_E_INVALID, "fine grids of %zu weight vectors x %d cells exceed the LDS", nw, sh.nf);
    sh.nsrc = (int)nsrc;
    sh.nw = (int)nw;
    sh.pos_cols = (int)pos_cols;
    hipStream_t st = (hipStream_t)stream;
    int rc = MRS_OK;
    double* dcorr = nullptr;
    NTRY(hipMalloc(&dcorr, n_out * sizeof(double)));
    NTRY(hipMemcpyAsync(dcorr, corr.data(), n_out * sizeof(double), hipMemcpyHostToDevice, st));
    NTRY(launch_nufft1(pos, reinterpret_cast<const double2*>(weights), dcorr, reinterpret_cast<double2*>(out),
                       (int)ncols, sh, st));
    NTRY(hipStreamSynchronize(st));
done:
    hipFree(dcorr);
    return rc;
}

extern "C" int mrs_build_coil_kspaces_nufft(const mrs_slice_input* inp, const mrs_acquisition* acq, mrs_c64* out,
                                            char* err, size_t errlen) {
    if (!inp || !acq || !out) return nfail(err, errlen, MRS_E_INVALID, "null argument");
    const size_t snx = inp->sim[0], sny = inp->sim[1], nx = inp->acq_matrix[0], ny = inp->acq_matrix[1];
    if (nx == 0 || ny == 0 || snx % nx || sny % ny)
        return nfail(err, errlen, MRS_E_INVALID, "sim grid must be an integer multiple of the acquired matrix");
    const size_t Q = inp->n_compartments;
    if (Q == 0 || Q > MRS_MAX_COMP) return nfail(err, errlen, MRS_E_INVALID, "1..%d compartments", MRS_MAX_COMP);
    // the NUFFT y-stage's domain (src/kspace.rs:775-792): uniform relaxation, no eddy polynomial,
    // ny >= 3, t affine in the line index per parity with the parity's polarity
    for (size_t c = 0; c < Q; c++)
        if (inp->t2[c].map || (inp->t_inhom && inp->t_inhom[c].map))
            return nfail(err, errlen, MRS_E_INVALID, "the NUFFT y-stage needs uniform relaxation (a T2 or T2' map was given)");
    if (acq->eddy_strength != 0.0 && inp->has_eddy_drive)
        return nfail(err, errlen, MRS_E_INVALID, "the NUFFT y-stage does not take the eddy polynomial (non-affine time profile)");
    if (ny < 3) return nfail(err, errlen, MRS_E_INVALID, "the NUFFT y-stage needs at least 3 phase-encode lines");
    if (2 * sny * 2 * Q * sizeof(double2) > kMaxNufftLds)
        return nfail(err, errlen, MRS_E_INVALID, "sim grid rows %zu too many for the NUFFT fine grids in LDS", sny);
    std::vector<double> t(ny), trf(ny), tread(ny);
    std::vector<int8_t> pol(ny);
    int rc = mrs_line_timing_for_acquisition(acq, nx, ny, t.data(), trf.data(), tread.data(), pol.data());
    if (rc) return nfail(err, errlen, rc, "line timing refused the acquisition");
    const double tau = (t[2] - t[0]) / 2.0, t0 = t[0], delta = t[1] - (tau + t0);
    double tmax = 1e-300;
    for (double v : t) tmax = std::max(tmax, std::fabs(v));
    for (size_t k = 0; k < ny; k++) {
        const double pred = tau * (double)k + t0 + ((k % 2 == 1) ? delta : 0.0);
        if (!(std::fabs(t[k] - pred) <= 1e-9 * tmax) || pol[k] != ((k % 2 == 1) ? -1 : 1))
            return nfail(err, errlen, MRS_E_INVALID, "the NUFFT y-stage needs line times affine per parity (line %zu)", k);
    }
    const size_t C = std::max<uint64_t>(acq->n_coils, 1);
    const size_t xs = nx / 2, ys = ny / 2, ox = snx / nx, oy = sny / ny;
    NufftSlice s{};
    s.snx = (int)snx; s.sny = (int)sny; s.nx = (int)nx; s.ny = (int)ny; s.Q = (int)Q; s.C = (int)C;
    s.ox = (int)ox; s.oy = (int)oy; s.sxs = (int)(ox * xs); s.sys = (int)(oy * ys);
    s.xoff = ((double)ox - 1.0) / 2.0; s.yoff = ((double)oy - 1.0) / 2.0;
    s.zc_c = (double)inp->z - ((double)inp->nz - 1.0) / 2.0;
    s.tau_s = tau / 1000.0; s.delta_s = delta / 1000.0; s.t_centre = (tau * (double)ys + t0) / 1000.0;
    s.signal_scale = acq->signal_scale; s.t_echo_s = acq->t_echo / 1000.0; s.eddy_phase = acq->eddy_phase;
    for (int i = 0; i < 3; i++) s.grad[i] = inp->has_eddy_drive ? inp->eddy_drive[i] : 0.0;
    s.do_dist = acq->do_distortions ? 1 : 0;
    s.gradient_echo = acq->echo == MRS_ECHO_GRADIENT ? 1 : 0;
    s.do_eddy_phase = (acq->eddy_phase != 0.0 && inp->has_eddy_drive) ? 1 : 0;
    // host tables: acquired rows, per-line uniform relaxation, x-stage twiddles, ghost table per line
    std::vector<uint8_t> mask(ny);
    mrs_sampling_mask(1, ny, acq, mask.data());
    std::vector<int32_t> arows;
    for (size_t k = 0; k < ny; k++) if (mask[k]) arows.push_back((int32_t)k);
    std::vector<double> rel(ny * Q, 1.0);
    for (size_t k = 0; k < ny; k++)
        for (size_t c = 0; c < Q; c++) {
            if (!acq->do_relaxation) continue;
            const double t2c = (double)inp->t2[c].value;
            const double tic = inp->t_inhom ? (double)inp->t_inhom[c].value : acq->t_inhom;
            rel[k * Q + c] = acq->echo == MRS_ECHO_SPIN ? std::exp(-trf[k] / t2c - std::fabs(t[k] / 1000.0) * 1000.0 / tic)
                                                       : std::exp(-trf[k] / t2c - trf[k] / tic);
        }
    std::vector<double2> tw(2 * nx * snx);
    const double cx = (double)(ox * xs) + s.xoff;
    for (int p = 0; p < 2; p++) {
        const double gh = p == 0 ? acq->ghost_offset : -acq->ghost_offset;
        for (size_t kx = 0; kx < nx; kx++) {
            const double kxn = ((double)kx - (double)xs + gh) / (double)snx;
            for (size_t x = 0; x < snx; x++) {
                const double ph = kTauN * kxn * ((double)x - cx);
                tw[((size_t)p * nx + kx) * snx + x] = make_double2(std::cos(ph), std::sin(ph));
            }
        }
    }
    std::vector<int8_t> table(ny);
    for (size_t k = 0; k < ny; k++) {
        const double shift = pol[k] < 0 ? -acq->ghost_offset : acq->ghost_offset;
        table[k] = shift < 0.0 ? 1 : 0;
    }
    NufftShape sh{};
    std::vector<double> corr;
    nufft_setup(sny, -(long)ys, ny, 1e-13, sh, corr);
    sh.nsrc = (int)sny;
    sh.nw = (int)(2 * Q);
    sh.pos_cols = (int)snx;
    const size_t P = snx * sny, nw = 2 * Q, ncols = C * snx;
    std::vector<float> comps(Q * P);
    for (size_t c = 0; c < Q; c++) std::copy(inp->compartments[c], inp->compartments[c] + P, comps.begin() + c * P);
    std::vector<double2> hk(C * nx * ny);
    float *dcomp = nullptr, *dfmap = nullptr;
    double *dph = nullptr, *dpos = nullptr, *dcorr = nullptr, *drel = nullptr;
    double2 *dw = nullptr, *drows = nullptr, *dtw = nullptr, *dk = nullptr;
    int8_t* dtab = nullptr;
    uint8_t* dacq = nullptr;
    int32_t* drw = nullptr;
    hipStream_t st = nullptr;
    NTRY(hipMalloc(&dcomp, Q * P * sizeof(float)));
    NTRY(hipMalloc(&dfmap, P * sizeof(float)));
    if (inp->phase0) NTRY(hipMalloc(&dph, P * sizeof(double)));
    NTRY(hipMalloc(&dpos, P * sizeof(double)));
    NTRY(hipMalloc(&dw, ncols * nw * sny * sizeof(double2)));
    NTRY(hipMalloc(&drows, ncols * nw * ny * sizeof(double2)));
    NTRY(hipMalloc(&dcorr, ny * sizeof(double)));
    NTRY(hipMalloc(&drel, ny * Q * sizeof(double)));
    NTRY(hipMalloc(&dtw, tw.size() * sizeof(double2)));
    NTRY(hipMalloc(&dk, hk.size() * sizeof(double2)));
    NTRY(hipMalloc(&dtab, ny));
    NTRY(hipMalloc(&dacq, ny));
    NTRY(hipMalloc(&drw, std::max<size_t>(arows.size(), 1) * 4));
    NTRY(hipMemcpy(dcomp, comps.data(), Q * P * sizeof(float), hipMemcpyHostToDevice));
    NTRY(hipMemcpy(dfmap, inp->fmap, P * sizeof(float), hipMemcpyHostToDevice));
    if (dph) NTRY(hipMemcpy(dph, inp->phase0, P * sizeof(double), hipMemcpyHostToDevice));
    NTRY(hipMemcpy(dcorr, corr.data(), ny * sizeof(double), hipMemcpyHostToDevice));
    NTRY(hipMemcpy(drel, rel.data(), ny * Q * sizeof(double), hipMemcpyHostToDevice));
    NTRY(hipMemcpy(dtw, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
    NTRY(hipMemcpy(dtab, table.data(), ny, hipMemcpyHostToDevice));
    NTRY(hipMemcpy(dacq, mask.data(), ny, hipMemcpyHostToDevice));
    if (!arows.empty()) NTRY(hipMemcpy(drw, arows.data(), arows.size() * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(nufft_sources, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, dcomp, dfmap, dph, s, dpos, dw);
    NTRY(hipGetLastError());
    NTRY(launch_nufft1(dpos, dw, dcorr, drows, (int)ncols, sh, st));
    hipLaunchKernelGGL(nufft_lines, dim3((unsigned)ny, (unsigned)C), dim3(256), snx * sizeof(double2), st, drows, drel,
                       dtw, dtab, dacq, s, 1.0 / (double)P, dk);
    NTRY(hipGetLastError());
    {
        const double sigma = acq->noise_variance > 0.0 ? std::sqrt(acq->noise_variance / (double)(nx * ny)) : 0.0;
        if (acq->n_spikes > 0 || sigma > 0.0) {
            hipLaunchKernelGGL(nufft_noise, dim3((unsigned)C), dim3(256), 0, st, dk, drw, (int)arows.size(), (int)nx,
                               (int)ny, inp->slice_seed, (int)acq->n_spikes, acq->spike_amplitude, sigma);
            NTRY(hipGetLastError());
        }
    }
    NTRY(hipMemcpy(hk.data(), dk, hk.size() * sizeof(double2), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < hk.size(); i++) out[i] = mrs_c64{hk[i].x, hk[i].y};
done:
    hipFree(dcomp); hipFree(dfmap); hipFree(dph); hipFree(dpos); hipFree(dw); hipFree(drows); hipFree(dcorr);
    hipFree(drel); hipFree(dtw); hipFree(dk); hipFree(dtab); hipFree(dacq); hipFree(drw);
    return rc;
}
