// recon.cpp -- host side of the batched reconstruction (include/mrsim_gpu.h): the window, the centred
// inverse 2D DFT and the Roemer combine of reconstruct_coils_full (src/kspace.rs:1301-1378), on the
// GPU for many units of one geometry.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "internal.h"

namespace mrs {
hipError_t launch_recon(const ReconGeom* g, const ReconItem* items, int nitems, int blocks_per_unit, int nty,
                        size_t lds, hipStream_t st);
size_t recon_lds_bytes(int nxP, int nyP);
hipError_t set_recon_lds_limit(size_t bytes);
hipError_t launch_series_epilogue(const float2* img, int nv, int nvox, const float* sigma, uint64_t seed, int mode,
                                  float* o0, float* o1, hipStream_t st);
}  // namespace mrs

using namespace mrs;

namespace {

constexpr double kTau = 6.283185307179586476925286766559;

int fail(char* err, size_t errlen, int code, const char* fmt, ...) {
    if (err && errlen) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(err, errlen, fmt, ap);
        va_end(ap);
    }
    return code;
}

// KspaceWindow::at (src/kspace.rs:140-161): radial windows, r = 2 hypot(kx, ky).
double window_at(int kind, double a_, double b_, double kx, double ky) {
    const double r = 2.0 * std::sqrt(kx * kx + ky * ky);
    switch (kind) {
    case MRS_WINDOW_NONE: return 1.0;
    case MRS_WINDOW_HANN: return r >= 1.0 ? 0.0 : 0.5 * (1.0 + std::cos(M_PI * r));
    case MRS_WINDOW_TUKEY: {
        const double a = std::min(std::max(a_, 0.0), 1.0);
        if (r <= 1.0 - a) return 1.0;
        if (r >= 1.0) return 0.0;
        return 0.5 * (1.0 + std::cos(M_PI * (r - (1.0 - a)) / std::max(a, 1e-12)));
    }
    case MRS_WINDOW_FERMI: return 1.0 / (1.0 + std::exp((r - a_) / std::max(b_, 1e-12)));
    default: return 1.0;
    }
}

uint16_t bf16_rne(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

// (re, im) -> packed bf16 hi and the bf16 of the residuals, as the kernels split their operands.
void split_pack(double re, double im, uint32_t& hi, uint32_t& lo) {
    const float fr = (float)re, fi = (float)im;
    const uint16_t hr = bf16_rne(fr), hm = bf16_rne(fi);
    uint32_t ur = (uint32_t)hr << 16, um = (uint32_t)hm << 16;
    float br, bm;
    std::memcpy(&br, &ur, 4);
    std::memcpy(&bm, &um, 4);
    hi = (uint32_t)hr | ((uint32_t)hm << 16);
    lo = (uint32_t)bf16_rne(fr - br) | ((uint32_t)bf16_rne(fi - bm) << 16);
}

// The centred inverse DFT kernel cis(-2 pi (k - n/2)(r - n/2)/n) over [r][k], zero-padded to nP,
// its angle reduced exactly in integers.
void dft_table(size_t n, size_t nP, std::vector<uint32_t>& hi, std::vector<uint32_t>& lo) {
    hi.assign(nP * nP, 0);
    lo.assign(nP * nP, 0);
    const long h = (long)(n / 2);
    for (size_t r = 0; r < n; r++)
        for (size_t k = 0; k < n; k++) {
            long m = (((long)k - h) * ((long)r - h)) % (long)n;
            if (m < 0) m += (long)n;
            const double th = -kTau * (double)m / (double)n;
            split_pack(std::cos(th), std::sin(th), hi[r * nP + k], lo[r * nP + k]);
        }
}

}  // namespace

struct mrs_recon {
    ReconGeom g{};
    ReconGeom* d_g = nullptr;
    std::vector<void*> allocs;
    ReconItem* d_items = nullptr;
    size_t cap = 0, nitems = 0, lds = 0;
    int nstrips = 1;
};

static void recon_free(mrs_recon* r) {
    for (void* a : r->allocs) hipFree(a);
    r->allocs.clear();
    if (r->d_items) hipFree(r->d_items);
    r->d_items = nullptr;
}

template <class T>
static int upload(mrs_recon* r, const T** dst, const std::vector<T>& v, char* err, size_t errlen) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T));
    if (e != hipSuccess) return fail(err, errlen, MRS_E_NOMEM, "hipMalloc: %s", hipGetErrorString(e));
    r->allocs.push_back(p);
    if (!v.empty() && (e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(err, errlen, MRS_E_DEVICE, "hipMemcpy: %s", hipGetErrorString(e));
    *dst = (const T*)p;
    return MRS_OK;
}

extern "C" int mrs_recon_create(size_t nx, size_t ny, size_t n_coils, const mrs_acquisition* acq,
                                const float* sensitivities, mrs_recon** out, char* err, size_t errlen) {
    if (!acq || !out) return fail(err, errlen, MRS_E_INVALID, "null argument");
    *out = nullptr;
    const size_t ncoils = std::max<uint64_t>(acq->n_coils, 1);
    if (n_coils != ncoils) return fail(err, errlen, MRS_E_INVALID, "one k-space per coil");
    if (nx < 1 || ny < 1) return fail(err, errlen, MRS_E_INVALID, "empty acquired matrix");
    if (nx > 128 || ny > 128)
        return fail(err, errlen, MRS_E_UNSUPPORTED, "reconstruction supports nx, ny <= 128 (got %zux%zu)", nx, ny);
    mrs_recon* r = new mrs_recon();
    const size_t nxP = (nx + 15) & ~size_t(15), nyP = (ny + 15) & ~size_t(15);
    ReconGeom& g = r->g;
    g.nx = (int)nx;
    g.ny = (int)ny;
    g.nxP = (int)nxP;
    g.nyP = (int)nyP;
    g.C = (int)ncoils;
    std::vector<uint32_t> fyh, fyl, fxh, fxl;
    dft_table(ny, nyP, fyh, fyl);
    dft_table(nx, nxP, fxh, fxl);
    // Roemer sensitivities on the acquired grid (real), and 1 / max(sum s^2, 1e-12)
    std::vector<float> sens(ncoils * nyP * nxP, 0.0f), rssum(nyP * nxP, 0.0f);
    std::vector<double> ssum(nyP * nxP, 0.0);
    for (size_t c = 0; c < ncoils; c++)
        for (size_t y = 0; y < ny; y++)
            for (size_t x = 0; x < nx; x++) {
                const double s = sensitivities ? (double)sensitivities[(c * ny + y) * nx + x]
                                               : ring_sensitivity(c, ncoils, (double)x, (double)y, nx, ny);
                sens[(c * nyP + y) * nxP + x] = (float)s;
                ssum[y * nxP + x] += s * s;
            }
    for (size_t i = 0; i < ssum.size(); i++) rssum[i] = (float)(1.0 / std::max(ssum[i], 1e-12));
    std::vector<float> win;
    if (acq->window != MRS_WINDOW_NONE) {
        win.resize(nx * ny);
        const double xs = (double)(nx / 2), ys = (double)(ny / 2);
        for (size_t ky = 0; ky < ny; ky++)
            for (size_t kx = 0; kx < nx; kx++)
                win[ky * nx + kx] = (float)window_at(acq->window, acq->window_a, acq->window_b,
                                                     ((double)kx - xs) / (double)nx, ((double)ky - ys) / (double)ny);
    }
    int rc = upload(r, &g.fy_hi, fyh, err, errlen);
    if (!rc) rc = upload(r, &g.fy_lo, fyl, err, errlen);
    if (!rc) rc = upload(r, &g.fx_hi, fxh, err, errlen);
    if (!rc) rc = upload(r, &g.fx_lo, fxl, err, errlen);
    if (!rc) rc = upload(r, &g.sens, sens, err, errlen);
    if (!rc) rc = upload(r, &g.rssum, rssum, err, errlen);
    if (!rc && !win.empty()) rc = upload(r, &g.window, win, err, errlen);
    if (!rc) {
        std::vector<ReconGeom> gv{g};
        const ReconGeom* dg = nullptr;
        rc = upload(r, &dg, gv, err, errlen);
        r->d_g = const_cast<ReconGeom*>(dg);
    }
    if (rc) {
        recon_free(r);
        delete r;
        return rc;
    }
    r->nstrips = (int)(nxP / 16);   // blocks per unit: one per 16-column x tile
    r->lds = recon_lds_bytes((int)nxP, (int)nyP);
    hipError_t e = set_recon_lds_limit(160 * 1024);
    if (e != hipSuccess) {
        recon_free(r);
        delete r;
        return fail(err, errlen, MRS_E_DEVICE, "cannot raise the reconstruction LDS limit: %s", hipGetErrorString(e));
    }
    *out = r;
    return MRS_OK;
}

// Upload the unit list (device pointers) once; mrs_recon_launch then runs it asynchronously.
extern "C" int mrs_recon_set_units(mrs_recon* r, const mrs_c32* const* kspace, mrs_c32* const* image, size_t n,
                                   char* err, size_t errlen) {
    if (!r) return fail(err, errlen, MRS_E_INVALID, "null reconstruction");
    if (n && (!kspace || !image)) return fail(err, errlen, MRS_E_INVALID, "null buffer list");
    std::vector<ReconItem> items(n);
    for (size_t i = 0; i < n; i++) {
        if (!kspace[i] || !image[i]) return fail(err, errlen, MRS_E_INVALID, "unit %zu: null buffer", i);
        items[i].kspace = reinterpret_cast<const float2*>(kspace[i]);
        items[i].image = reinterpret_cast<float2*>(image[i]);
    }
    hipError_t e;
    if ((e = hipDeviceSynchronize()) != hipSuccess)
        return fail(err, errlen, MRS_E_DEVICE, "hipDeviceSynchronize: %s", hipGetErrorString(e));
    if (n > r->cap) {
        if (r->d_items) hipFree(r->d_items);
        r->d_items = nullptr;
        r->cap = 0;
        if ((e = hipMalloc(&r->d_items, n * sizeof(ReconItem))) != hipSuccess)
            return fail(err, errlen, MRS_E_NOMEM, "hipMalloc: %s", hipGetErrorString(e));
        r->cap = n;
    }
    if (n && (e = hipMemcpy(r->d_items, items.data(), n * sizeof(ReconItem), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(err, errlen, MRS_E_DEVICE, "hipMemcpy: %s", hipGetErrorString(e));
    r->nitems = n;
    return MRS_OK;
}

// Asynchronous on `stream`; no allocation or copy, so it can be recorded into a HIP graph.
extern "C" int mrs_recon_launch(mrs_recon* r, void* stream, char* err, size_t errlen) {
    if (!r) return fail(err, errlen, MRS_E_INVALID, "null reconstruction");
    hipError_t e = launch_recon(r->d_g, r->d_items, (int)r->nitems, r->nstrips, r->g.nyP / 16, r->lds,
                                (hipStream_t)stream);
    if (e != hipSuccess) return fail(err, errlen, MRS_E_DEVICE, "reconstruction: %s", hipGetErrorString(e));
    return MRS_OK;
}

extern "C" int mrs_recon_run(mrs_recon* r, const mrs_c32* const* kspace, mrs_c32* const* image, size_t n, void* stream,
                             char* err, size_t errlen) {
    int rc = mrs_recon_set_units(r, kspace, image, n, err, errlen);
    if (rc || n == 0) return rc;
    if ((rc = mrs_recon_launch(r, stream, err, errlen))) return rc;
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return fail(err, errlen, MRS_E_DEVICE, "reconstruction: %s", hipGetErrorString(e));
    return MRS_OK;
}

extern "C" int mrs_recon_destroy(mrs_recon* r) {
    if (!r) return MRS_OK;
    hipDeviceSynchronize();
    recon_free(r);
    delete r;
    return MRS_OK;
}

// reconstruct_coils (src/kspace.rs:1290-1298) for host buffers: coil k-spaces (n_coils x nx*ny,
// kx + nx*ky, complex f64) -> the combined image (x + nx*y, complex f64); GRAPPA first under
// acceleration.
extern "C" int mrs_reconstruct_coils(const mrs_c64* coil_kspace, size_t n_coils, size_t nx, size_t ny,
                                     const mrs_acquisition* acq, const mrs_c64* grappa_weights, mrs_c64* out,
                                     char* err, size_t errlen) {
    if (!coil_kspace || !out || !acq) return fail(err, errlen, MRS_E_INVALID, "null argument");
    if (acq->accel > 1 && !grappa_weights)
        return fail(err, errlen, MRS_E_INVALID, "an accelerated slice needs its GRAPPA kernel (fitted on a reference scan)");
    mrs_recon* r = nullptr;
    int rc = mrs_recon_create(nx, ny, n_coils, acq, nullptr, &r, err, errlen);
    if (rc) return rc;
    const size_t nk = n_coils * nx * ny;
    const size_t nw = acq->accel > 1 ? (acq->accel - 1) * n_coils * 6 * n_coils : 0;
    std::vector<float2> hk(nk), hi(nx * ny);
    for (size_t i = 0; i < nk; i++) hk[i] = make_float2((float)coil_kspace[i].re, (float)coil_kspace[i].im);
    float2 *dk = nullptr, *di = nullptr;
    mrs_c64* dw = nullptr;
    hipError_t e = hipMalloc(&dk, nk * sizeof(float2));
    if (e == hipSuccess) e = hipMalloc(&di, nx * ny * sizeof(float2));
    if (e == hipSuccess) e = hipMemcpy(dk, hk.data(), nk * sizeof(float2), hipMemcpyHostToDevice);
    if (e == hipSuccess && nw) e = hipMalloc(&dw, nw * sizeof(mrs_c64));
    if (e == hipSuccess && nw) e = hipMemcpy(dw, grappa_weights, nw * sizeof(mrs_c64), hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = fail(err, errlen, MRS_E_DEVICE, "device buffers: %s", hipGetErrorString(e));
    mrs_c32* kp = reinterpret_cast<mrs_c32*>(dk);
    if (!rc && nw) {
        std::vector<uint8_t> mask(ny);
        mrs_sampling_mask(1, ny, acq, mask.data());
        rc = mrs_grappa_apply(&kp, 1, n_coils, nx, ny, acq->accel, mask.data(), dw, nullptr, err, errlen);
    }
    if (!rc) {
        const mrs_c32* kc = kp;
        mrs_c32* ip = reinterpret_cast<mrs_c32*>(di);
        rc = mrs_recon_run(r, &kc, &ip, 1, nullptr, err, errlen);
    }
    if (!rc && (e = hipMemcpy(hi.data(), di, nx * ny * sizeof(float2), hipMemcpyDeviceToHost)) != hipSuccess)
        rc = fail(err, errlen, MRS_E_DEVICE, "copy back: %s", hipGetErrorString(e));
    if (!rc)
        for (size_t i = 0; i < nx * ny; i++) out[i] = mrs_c64{(double)hi[i].x, (double)hi[i].y};
    hipFree(dk);
    hipFree(di);
    hipFree(dw);
    mrs_recon_destroy(r);
    return rc;
}

// simulate_slice (src/kspace.rs:1229-1235, 1273-1284) on the GPU: coil k-spaces from the drop-in
// forward, the GRAPPA kernel fitted on the slice's own reference scan under acceleration
// (grappa_reference_kernel, :1256-1262, on the seed salted by SEED_SALT_GRAPPA_CAL), the
// reconstruction, and the f32 cast of the combined image.
extern "C" int mrs_simulate_slice(const mrs_slice_input* inp, const mrs_acquisition* acq, float* out, char* err,
                                  size_t errlen) {
    if (!inp || !acq || !out) return fail(err, errlen, MRS_E_INVALID, "null argument");
    const size_t nx = inp->acq_matrix[0], ny = inp->acq_matrix[1];
    const size_t nc = std::max<uint64_t>(acq->n_coils, 1);
    const size_t nk = nc * nx * ny;
    std::vector<mrs_c64> coils(nk), img(nx * ny);
    std::vector<mrs_c64> w;
    int rc = MRS_OK;
    if (acq->accel > 1) {
        mrs_slice_input ref = *inp;
        ref.slice_seed = inp->slice_seed ^ 0x47525043414CULL;
        ref.has_eddy_drive = 0;
        ref.has_prep_drive = 0;
        mrs_acquisition acq_ref = *acq;
        acq_ref.accel = 1;
        acq_ref.n_spikes = 0;
        rc = mrs_build_coil_kspaces(&ref, &acq_ref, nullptr, nullptr, coils.data(), err, errlen);
        if (rc) return rc;
        std::vector<float2> hk(nk);
        for (size_t i = 0; i < nk; i++) hk[i] = make_float2((float)coils[i].re, (float)coils[i].im);
        std::vector<uint8_t> lines(ny);
        mrs_sampling_mask(1, ny, &acq_ref, lines.data());
        const size_t nw = (acq->accel - 1) * nc * 6 * nc;
        w.resize(nw);
        float2* dk = nullptr;
        mrs_c64* dw = nullptr;
        hipError_t e = hipMalloc(&dk, nk * sizeof(float2));
        if (e == hipSuccess) e = hipMalloc(&dw, nw * sizeof(mrs_c64));
        if (e == hipSuccess) e = hipMemcpy(dk, hk.data(), nk * sizeof(float2), hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(err, errlen, MRS_E_DEVICE, "device buffers: %s", hipGetErrorString(e));
        if (!rc)
            rc = mrs_grappa_fit(reinterpret_cast<const mrs_c32*>(dk), nc, nx, ny, acq->accel, acq->acs_lines,
                                lines.data(), dw, nullptr, err, errlen);
        if (!rc && (e = hipMemcpy(w.data(), dw, nw * sizeof(mrs_c64), hipMemcpyDeviceToHost)) != hipSuccess)
            rc = fail(err, errlen, MRS_E_DEVICE, "copy back: %s", hipGetErrorString(e));
        hipFree(dk);
        hipFree(dw);
        if (rc) return rc;
    }
    rc = mrs_build_coil_kspaces(inp, acq, nullptr, nullptr, coils.data(), err, errlen);
    if (!rc) rc = mrs_reconstruct_coils(coils.data(), nc, nx, ny, acq, w.empty() ? nullptr : w.data(), img.data(), err,
                                        errlen);
    if (rc) return rc;
    for (size_t i = 0; i < nx * ny; i++) {
        out[2 * i] = (float)img[i].re;
        out[2 * i + 1] = (float)img[i].im;
    }
    return MRS_OK;
}

// The tail of acquire_volumes (src/kspace.rs:2036-2060, 1840-1846) over a series' combined images.
extern "C" int mrs_series_epilogue(const mrs_c32* images, size_t n_volumes, size_t nvox, const float* noise_sigma,
                                   uint64_t receiver_seed, int32_t mode, float* out0, float* out1, void* stream,
                                   char* err, size_t errlen) {
    if (!images || !out0 || !out1) return fail(err, errlen, MRS_E_INVALID, "null argument");
    if (mode != 0 && mode != 1) return fail(err, errlen, MRS_E_INVALID, "mode must be 0 (re, im) or 1 (magnitude, phase)");
    if (n_volumes * nvox > (size_t)INT32_MAX) return fail(err, errlen, MRS_E_INVALID, "series too large for one launch");
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = launch_series_epilogue(reinterpret_cast<const float2*>(images), (int)n_volumes, (int)nvox, noise_sigma,
                                          receiver_seed, mode, out0, out1, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(err, errlen, MRS_E_DEVICE, "series epilogue: %s", hipGetErrorString(e));
    return MRS_OK;
}
