// tables.cpp -- host-side protocol tables of the forward (no HIP): the EPI timing table, the acquired
// PE rows, the ring coil sensitivities, the bf16 operand split of host-built tables, and the 3D
// gradient-echo block timing, line table and timing checks, with their C entry points
// (include/mrsim_gpu.h).  Plain C++17, so the same source builds under the host sanitizers.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "tables.h"

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

}  // namespace

namespace mrs {


// SingleShotEpi (src/readout.rs:21-69): tick -> (kx, ky) snake trajectory and its times.
struct Epi {
    size_t nx, ny;
    double t_line, t_echo;
    bool reverse;
    double dt() const { return t_line / (double)nx; }
    double half() const { return (double)(nx * ny) * dt() / 2.0; }
    void index(size_t tick, size_t& kx, size_t& ky) const {
        size_t x = tick % nx, y = tick / nx;
        if (!reverse) {
            y = ny - 1 - y;
            if (y % 2 == 1) x = nx - x - 1;
        } else if (y % 2 == 1) {
            x = nx - x - 1;
        }
        kx = x;
        ky = y;
    }
    double from_echo(size_t tick) const { return dt() * ((double)tick + 0.5) - half(); }
    double from_prep(size_t tick) const { return (double)tick * dt() + dt() / 2.0; }
    double from_rf(size_t tick) const { return t_echo + from_echo(tick); }
};

size_t pf_skipped_lines(size_t ny, const mrs_acquisition& a) {
    if (a.partial_fourier < 1.0 && a.pf_mode == MRS_PF_SCANNER) {
        size_t keep = std::min((size_t)std::round((double)ny * a.partial_fourier), ny);
        return ny - keep;
    }
    return 0;
}

void timing_table(const mrs_acquisition& a, size_t nx, size_t ny, double* t, double* trf, double* tread, int8_t* pol) {
    Epi e{nx, ny, a.t_line, a.t_echo, a.reverse_phase != 0};
    std::fill(t, t + ny, 0.0);
    std::fill(trf, trf + ny, 0.0);
    std::fill(tread, tread + ny, 0.0);
    const size_t centre = nx / 2;
    for (size_t tick = 0; tick < nx * ny; tick++) {
        size_t kx, ky;
        e.index(tick, kx, ky);
        if (kx != centre) continue;
        t[ky] = e.from_echo(tick);
        trf[ky] = e.from_rf(tick);
        tread[ky] = e.from_prep(tick);
    }
    for (size_t k = 0; k < ny; k++) pol[k] = (k % 2 == 1) ? -1 : 1;
    const double offset = (double)pf_skipped_lines(ny, a) * a.t_line;
    if (offset > 0.0)
        for (size_t k = 0; k < ny; k++) tread[k] -= offset;
}

// Acquired PE rows (sampling_mask rows, src/kspace.rs:546-588).
std::vector<uint8_t> acquired_rows(size_t ny, const mrs_acquisition& a) {
    std::vector<uint8_t> rows(ny, 0);
    const size_t ys = ny / 2;
    const size_t accel = std::max<uint64_t>(a.accel, 1);
    const size_t keep_n = (size_t)std::round((double)ny * a.partial_fourier);
    const bool pf = a.partial_fourier < 1.0;
    for (size_t k = 0; k < ny; k++) {
        bool skip = false;
        if (pf && a.pf_mode == MRS_PF_SCANNER) {
            skip = a.reverse_phase ? (k < ny - keep_n) : (k >= keep_n);
        } else if (pf && a.pf_mode == MRS_PF_CONTIGUOUS) {
            skip = a.reverse_phase ? (k >= keep_n) : (k < ny - keep_n);
        } else if (pf) {
            skip = a.reverse_phase ? ((double)k > std::ceil((double)ny * a.partial_fourier))
                                   : (((double)k < std::floor((double)ny * (1.0 - a.partial_fourier))) &&
                                      (k > 0 || ny % 2 == 1));
        }
        if (skip) continue;
        if (accel <= 1 || k % accel == ys % accel) rows[k] = 1;
    }
    return rows;
}


uint16_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float bf16_to_float(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
// FP64 complex -> packed bf16 hi and residual lo, via FP32 (the device's operand precision).
void split_c(double re, double im, uint32_t& hi, uint32_t& lo) {
    float fr = (float)re, fi = (float)im;
    uint16_t hr = bf16_rne(fr), hm = bf16_rne(fi);
    float rr = (float)(re - (double)bf16_to_float(hr)), ri = (float)(im - (double)bf16_to_float(hm));
    hi = (uint32_t)hr | ((uint32_t)hm << 16);
    lo = (uint32_t)bf16_rne(rr) | ((uint32_t)bf16_rne(ri) << 16);
}

// coil_sensitivity (src/kspace.rs:226-237): the real ring Gaussian plus 0.15, in acquired-voxel units.
double ring_sensitivity(size_t coil, size_t n_coils, double x, double y, size_t nx, size_t ny) {
    if (n_coils <= 1) return 1.0;
    const double cx = (double)nx / 2.0, cy = (double)ny / 2.0;
    const double m = (double)std::max(nx, ny);
    const double r = 0.6 * m;
    const double ang = kTau * (double)coil / (double)n_coils;
    const double px = cx + r * std::cos(ang), py = cy + r * std::sin(ang);
    const double dx = x - px, dy = y - py;
    const double sigma = 0.9 * m;
    return std::exp(-(dx * dx + dy * dy) / (2.0 * sigma * sigma)) + 0.15;
}
// readout::grase_block, src/readout.rs:206-221: line ky = sy + ky_segments*j sits at acquisition
// index a = j (reversed) or epi-1-j within its segment's block, t = (a - (epi-1)/2) t_line, and the
// readout polarity alternates with a.
int grase_block(size_t ny, size_t kseg, double t_line, bool reverse, std::vector<double>& t,
                std::vector<int8_t>& pol, char* err, size_t errlen) {
    if (kseg == 0 || ny % kseg != 0)
        return fail(err, errlen, MRS_E_INVALID, "%zu phase-encode lines do not divide into %zu ky segments", ny, kseg);
    const size_t epi = ny / kseg;
    t.assign(ny, 0.0);
    pol.assign(ny, 0);
    for (size_t ky = 0; ky < ny; ky++) {
        const size_t j = ky / kseg;
        const size_t a = reverse ? j : epi - 1 - j;
        t[ky] = ((double)a - ((double)epi - 1.0) / 2.0) * t_line;
        pol[ky] = (a % 2 == 1) ? -1 : 1;
    }
    return MRS_OK;
}

// readout::kz_order, src/readout.rs:170-189: linear, or centric (the centre nz/2, then +1, -1, +2, ...).
std::vector<size_t> kz_order(size_t nz, bool centric) {
    std::vector<size_t> out;
    if (!centric) {
        for (size_t p = 0; p < nz; p++) out.push_back(p);
        return out;
    }
    const size_t c = nz / 2;
    out.push_back(c);
    for (size_t d = 1; out.size() < nz; d++) {
        if (c + d < nz) out.push_back(c + d);
        if (d <= c && out.size() < nz) out.push_back(c - d);
    }
    return out;
}

}  // namespace mrs

using namespace mrs;

// ------------------------------------------------------------------ host tables (C ABI)

extern "C" int mrs_acquisition_default(mrs_acquisition* a) {
    if (!a) return MRS_E_INVALID;
    std::memset(a, 0, sizeof *a);
    a->t_line = 1.0;
    a->t_echo = 90.0;
    a->t_inhom = 50.0;
    a->signal_scale = 100.0;
    a->do_distortions = 1;
    a->do_relaxation = 1;
    a->partial_fourier = 1.0;
    a->pf_mode = MRS_PF_FIBERFOX;
    a->eddy_tau = 70.0;
    a->spike_amplitude = 1.0;
    a->window = MRS_WINDOW_NONE;
    a->n_coils = 1;
    a->accel = 1;
    a->acs_lines = 24;
    a->echo = MRS_ECHO_SPIN;
    return MRS_OK;
}

extern "C" int mrs_line_timing_for_acquisition(const mrs_acquisition* acq, size_t nx, size_t ny, double* t_ms,
                                               double* trf_ms, double* tread_ms, int8_t* polarity) {
    if (!acq || !t_ms || !trf_ms || !tread_ms || !polarity || nx == 0) return MRS_E_INVALID;
    timing_table(*acq, nx, ny, t_ms, trf_ms, tread_ms, polarity);
    return MRS_OK;
}

extern "C" int mrs_sampling_mask(size_t nx, size_t ny, const mrs_acquisition* acq, uint8_t* mask) {
    if (!acq || !mask) return MRS_E_INVALID;
    std::vector<uint8_t> rows = acquired_rows(ny, *acq);
    for (size_t k = 0; k < ny; k++)
        for (size_t x = 0; x < nx; x++) mask[x + nx * k] = rows[k];
    return MRS_OK;
}

extern "C" int mrs_ring_coil_maps(size_t coil0, size_t n, size_t ncoils, size_t snx, size_t sny, size_t nx, size_t ny,
                                  double scale, mrs_c32* out) {
    if (!out || nx == 0 || ny == 0 || snx % nx || sny % ny) return MRS_E_INVALID;
    const size_t ox = snx / nx, oy = sny / ny;
    const double xoff = ((double)ox - 1.0) / 2.0, yoff = ((double)oy - 1.0) / 2.0;
    for (size_t c = 0; c < n; c++)
        for (size_t y = 0; y < sny; y++)
            for (size_t x = 0; x < snx; x++) {
                const double xa = ((double)x - xoff) / (double)ox, ya = ((double)y - yoff) / (double)oy;
                const double s = scale * ring_sensitivity(coil0 + c, ncoils, xa, ya, nx, ny);
                out[c * snx * sny + x + snx * y] = mrs_c32{(float)s, 0.0f};
            }
    return MRS_OK;
}

extern "C" int mrs_grase_block(size_t ny, size_t ky_segments, double t_line_ms, int32_t reverse_phase, double* t_ms,
                               int8_t* polarity, char* err, size_t errlen) {
    std::vector<double> t;
    std::vector<int8_t> pol;
    int rc = grase_block(ny, ky_segments, t_line_ms, reverse_phase != 0, t, pol, err, errlen);
    if (rc) return rc;
    if (t_ms) std::copy(t.begin(), t.end(), t_ms);
    if (polarity) std::copy(pol.begin(), pol.end(), polarity);
    return MRS_OK;
}

// readout::ge3d_lines, src/readout.rs:428-462.
extern "C" int mrs_ge3d_lines(size_t ny, size_t nz, size_t ky_segments, double t_line_ms, int32_t reverse_phase,
                              size_t kz_segments, int32_t kz_order_kind, int32_t* shot, int32_t* excitation,
                              size_t* n_shots, size_t* n_exc, size_t* e_c, char* err, size_t errlen) {
    if (nz < 2)
        return fail(err, errlen, MRS_E_INVALID, "a 3D gradient-echo train on %zu partition is a 2D readout", nz);
    std::vector<double> t;
    std::vector<int8_t> pol;
    int rc = grase_block(ny, ky_segments, t_line_ms, reverse_phase != 0, t, pol, err, errlen);
    if (rc) return rc;
    if (kz_segments == 0 || nz % kz_segments != 0)
        return fail(err, errlen, MRS_E_INVALID, "%zu partitions do not divide into %zu kz segments", nz, kz_segments);
    const size_t nexc = nz / kz_segments, epi = ny / ky_segments;
    const std::vector<size_t> order = kz_order(nz, kz_order_kind != 0);
    std::vector<int32_t> sh(ny * nz, -1), ex(ny * nz, -1);
    for (size_t sy = 0; sy < ky_segments; sy++)
        for (size_t sz = 0; sz < kz_segments; sz++)
            for (size_t j = 0; j < nexc; j++) {
                const size_t p = order[sz + kz_segments * j];
                for (size_t i = 0; i < epi; i++) {
                    const size_t at = p * ny + sy + ky_segments * i;
                    if (sh[at] >= 0) return fail(err, errlen, MRS_E_INVALID, "line read twice");
                    sh[at] = (int32_t)(sy * kz_segments + sz);
                    ex[at] = (int32_t)j;
                }
            }
    if (shot) std::copy(sh.begin(), sh.end(), shot);
    if (excitation) std::copy(ex.begin(), ex.end(), excitation);
    if (n_shots) *n_shots = ky_segments * kz_segments;
    if (n_exc) *n_exc = nexc;
    if (e_c) *e_c = (size_t)(std::find(order.begin(), order.end(), nz / 2) - order.begin()) / kz_segments;
    return MRS_OK;
}

// readout::check_ge3d_timing, src/readout.rs:465-501: each echo's block spans TE_e +- half, half =
// max |t| + t_line/2.
extern "C" int mrs_check_ge3d_timing(const double* tes, size_t n_echoes, double exc_spacing_ms,
                                     double excitation_time_ms, size_t ny, size_t ky_segments, double t_line_ms,
                                     int32_t reverse_phase, size_t n_exc, double t_exc_ms, double tr_ms, char* err,
                                     size_t errlen) {
    if (n_echoes == 0 || !tes)
        return fail(err, errlen, MRS_E_INVALID, "a 3D gradient-echo train needs at least one echo time");
    std::vector<double> t;
    std::vector<int8_t> pol;
    int rc = grase_block(ny, ky_segments, t_line_ms, reverse_phase != 0, t, pol, err, errlen);
    if (rc) return rc;
    double half = 0.0;
    for (double v : t) half = std::max(half, std::fabs(v));
    half += t_line_ms / 2.0;
    const double pulse = excitation_time_ms / 2.0;
    const size_t epi = ny / ky_segments;
    if (tes[0] - half < pulse - 1e-9)
        return fail(err, errlen, MRS_E_INVALID,
                    "echo 1 at %.4f ms: its %zu-line block starts %.4f ms after the excitation, inside the excitation "
                    "pulse (half of %.3f ms); a later echo time or a shorter block would fit",
                    tes[0], epi, tes[0] - half, excitation_time_ms);
    for (size_t e = 1; e < n_echoes; e++)
        if (tes[e] - half < tes[e - 1] + half - 1e-9)
            return fail(err, errlen, MRS_E_INVALID,
                        "echo %zu at %.4f ms and echo %zu at %.4f ms: their blocks of %.4f ms each side overlap", e,
                        tes[e - 1], e + 1, tes[e], half);
    const double end = tes[n_echoes - 1] + half;
    if (n_exc > 1 && end > exc_spacing_ms - pulse + 1e-9)
        return fail(err, errlen, MRS_E_INVALID,
                    "the last echo's block ends %.4f ms after its excitation, after the next excitation pulse starts "
                    "at %.4f ms (excitation spacing %.4f ms, pulse %.3f ms): the spacing must be at least %.4f ms",
                    end, exc_spacing_ms - pulse, exc_spacing_ms, excitation_time_ms, end + pulse);
    const double last = t_exc_ms + ((double)n_exc - 1.0) * exc_spacing_ms + end;
    if (last > tr_ms + 1e-9)
        return fail(err, errlen, MRS_E_INVALID,
                    "the train ends at %.4f ms (%zu excitations %.4f ms apart from %.4f ms), after the %.4f ms "
                    "repetition",
                    last, n_exc, exc_spacing_ms, t_exc_ms, tr_ms);
    return MRS_OK;
}

