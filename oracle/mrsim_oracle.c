/*
 * mrsim_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A CPU FP64 restatement of the reference's exact k-space forward model and of the per-slice
 * reconstruction that its golden fixtures observe.  Only tests/, the smoke check and bench.py's
 * CPU-baseline leg may load this library, and only as the checker.  The product path never links
 * or calls it.
 *
 * Built with -ffp-contract=off (the reference is Rust, which never fuses multiply-adds) and linked
 * against glibc libm (what Rust std's f64::{exp,ln,sin,cos} call on linux-gnu), and written in the
 * reference's operation order, so that its results match the reference's golden fixtures
 * bit for bit.  Each function cites the reference lines (src/kspace.rs unless noted) it restates.
 *
 * Parity pins (tests/test_oracle.py):
 *   - the 39 per-slice golden fixtures tests/golden/p4_baseline/default/slice-*.bin
 *     (tests/golden_baseline.rs:28-89 of the reference), compared bit for bit;
 *   - the two series records series-{uniform,mapped}.bin (tests/golden_baseline.rs:91-124), through
 *     acquire_volumes with per-slice GRAPPA kernels and the hbcd_like object phase, bit for bit
 *     (tests/test_oracle_series.py);
 *   - EXPECTED_BITS_STD (src/kspace.rs:3484), the first eight raw coefficients, bit for bit;
 *   - the literal per-line sum (src/kspace.rs:2225-2406) within 1e-10 x peak;
 *   - the line-timing table against SingleShotEpi tick by tick (src/kspace.rs:3337-3353);
 *   - closed forms: gradient echo = spin echo x e^{i2pi f TE}, global phase rotation, linearity,
 *     unacquired rows zero, inverse DFT of a single k-space sample.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mrsim_gpu.h"

#define TAU 6.283185307179586476925286766559

typedef mrs_c64 C;

/* PhaseModel (src/phase.rs:129-137): global + BackgroundPhase coefficients (1, x, y, z, x^2, y^2,
 * z^2, xy, xz, yz) + an optional PrepPhase. */
typedef struct {
    double global;
    double coeffs[10];
    int32_t has_prep;
    double c_q, sigma_dx, sigma_rot;
} orc_phase_model;

static C c_add(C a, C b) { C r = {a.re + b.re, a.im + b.im}; return r; }
static C c_mul(C a, C b) { C r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static C c_scale(C a, double s) { C r = {a.re * s, a.im * s}; return r; }
static C c_cis(double th) { C r = {cos(th), sin(th)}; return r; }
static double c_abs(C a) { return sqrt(a.re * a.re + a.im * a.im); }

/* ---------------------------------------------------------------- Rng (:59-77) */
uint64_t orc_rng_next(uint64_t* s) {
    *s += 0x9E3779B97F4A7C15ULL;
    uint64_t z = *s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double orc_rng_unit(uint64_t* s) {
    return ((double)(orc_rng_next(s) >> 11) + 0.5) / (double)(1ULL << 53);
}
double orc_rng_gauss(uint64_t* s) {
    double u1 = orc_rng_unit(s);
    double u2 = orc_rng_unit(s);
    return sqrt(-2.0 * log(u1)) * cos(TAU * u2);
}

/* ---------------------------------------------------------------- Acquisition::default (:239-267) */
void orc_acquisition_default(mrs_acquisition* a) {
    memset(a, 0, sizeof *a);
    a->t_line = 1.0;
    a->t_echo = 90.0;
    a->t_inhom = 50.0;
    a->signal_scale = 100.0;
    a->reverse_phase = 0;
    a->do_distortions = 1;
    a->do_relaxation = 1;
    a->noise_variance = 0.0;
    a->partial_fourier = 1.0;
    a->pf_mode = MRS_PF_FIBERFOX;
    a->ghost_offset = 0.0;
    a->eddy_strength = 0.0;
    a->eddy_quad = 0.0;
    a->eddy_phase = 0.0;
    a->eddy_tau = 70.0;
    a->n_spikes = 0;
    a->spike_amplitude = 1.0;
    a->window = MRS_WINDOW_NONE;
    a->n_coils = 1;
    a->accel = 1;
    a->acs_lines = 24;
    a->seed = 0;
    a->echo = MRS_ECHO_SPIN;
}

/* ---------------------------------------------------------------- SingleShotEpi (src/readout.rs:33-69) */
typedef struct { size_t kx_max, ky_max; double t_line, t_echo; int reverse_phase; } Epi;
static double epi_dt(const Epi* e) { return e->t_line / (double)e->kx_max; }
static double epi_half(const Epi* e) { return (double)(e->kx_max * e->ky_max) * epi_dt(e) / 2.0; }
static void epi_index(const Epi* e, size_t tick, size_t* kx, size_t* ky) {
    size_t x = tick % e->kx_max, y = tick / e->kx_max;
    if (!e->reverse_phase) {
        y = e->ky_max - 1 - y;
        if (y % 2 == 1) x = e->kx_max - x - 1;
    } else if (y % 2 == 1) {
        x = e->kx_max - x - 1;
    }
    *kx = x;
    *ky = y;
}
static double epi_t_echo(const Epi* e, size_t tick) { return epi_dt(e) * ((double)tick + 0.5) - epi_half(e); }
static double epi_t_prep(const Epi* e, size_t tick) { return (double)tick * epi_dt(e) + epi_dt(e) / 2.0; }
static double epi_t_rf(const Epi* e, size_t tick) { return e->t_echo + epi_t_echo(e, tick); }

/* pf_skipped_lines (:110-116) */
static size_t pf_skipped(size_t ny, const mrs_acquisition* a) {
    if (a->partial_fourier < 1.0 && a->pf_mode == MRS_PF_SCANNER) {
        size_t keep = (size_t)round((double)ny * a->partial_fourier);
        if (keep > ny) keep = ny;
        return ny - keep;
    }
    return 0;
}

/* line_times (:271-284) + LineTiming::from_epi (:301-305) + for_acquisition (:310-326) */
void orc_line_timing(const mrs_acquisition* a, size_t nx, size_t ny, double* t, double* trf, double* tread,
                     int8_t* pol) {
    Epi e = {nx, ny, a->t_line, a->t_echo, a->reverse_phase};
    size_t xs = nx / 2;
    for (size_t k = 0; k < ny; k++) t[k] = trf[k] = tread[k] = 0.0;
    for (size_t tick = 0; tick < nx * ny; tick++) {
        size_t kx, ky;
        epi_index(&e, tick, &kx, &ky);
        if (kx == xs) {
            t[ky] = epi_t_echo(&e, tick);
            trf[ky] = epi_t_rf(&e, tick);
            tread[ky] = epi_t_prep(&e, tick);
        }
    }
    for (size_t k = 0; k < ny; k++) pol[k] = (k % 2 == 1) ? -1 : 1;
    double offset = (double)pf_skipped(ny, a) * a->t_line;
    if (offset > 0.0)
        for (size_t k = 0; k < ny; k++) tread[k] -= offset;
}

/* sampling_mask (:546-588) */
void orc_sampling_mask(size_t nx, size_t ny, const mrs_acquisition* a, uint8_t* m) {
    size_t ys = ny / 2;
    size_t accel = a->accel > 1 ? a->accel : 1;
    size_t keep_n = (size_t)round((double)ny * a->partial_fourier);
    memset(m, 0, nx * ny);
    for (size_t kyi = 0; kyi < ny; kyi++) {
        if (a->partial_fourier < 1.0 && a->pf_mode == MRS_PF_SCANNER) {
            int skip = a->reverse_phase ? (kyi < ny - keep_n) : (kyi >= keep_n);
            if (skip) continue;
        } else if (a->partial_fourier < 1.0 && a->pf_mode == MRS_PF_CONTIGUOUS) {
            int skip = a->reverse_phase ? (kyi >= keep_n) : (kyi < ny - keep_n);
            if (skip) continue;
        } else if (a->partial_fourier < 1.0) {
            int skip;
            if (a->reverse_phase)
                skip = (double)kyi > ceil((double)ny * a->partial_fourier);
            else
                skip = ((double)kyi < floor((double)ny * (1.0 - a->partial_fourier))) && (kyi > 0 || ny % 2 == 1);
            if (skip) continue;
        }
        int acquired = accel <= 1 || kyi % accel == ys % accel;
        if (!acquired) continue;
        for (size_t kxi = 0; kxi < nx; kxi++) m[kxi + nx * kyi] = 1;
    }
}

/* coil_sensitivity (:226-237) */
double orc_coil_sensitivity(size_t coil, size_t n_coils, double x, double y, size_t nx, size_t ny) {
    if (n_coils <= 1) return 1.0;
    double cx = (double)nx / 2.0, cy = (double)ny / 2.0;
    size_t mx = nx > ny ? nx : ny;
    double r = 0.6 * (double)mx;
    double ang = TAU * (double)coil / (double)n_coils;
    double px = cx + r * cos(ang), py = cy + r * sin(ang);
    double dx = x - px, dy = y - py;
    double d2 = dx * dx + dy * dy;
    double sigma = 0.9 * (double)mx;
    return exp(-d2 / (2.0 * sigma * sigma)) + 0.15;
}

/* T2Slice value at voxel i */
static double t2_at(const mrs_t2_slice* s, size_t i) { return s->map ? (double)s->map[i] : (double)s->value; }

/* ---------------------------------------------------------------- XStage (:1052-1140)
 * Default build: twiddle tables tw[pol][x + snx*kxi] (:1091-1101, :1126-1137).  With the FFT x-stage
 * switched on (the `kspace` feature's form, :1077-1088, :1110-1124): the input ghost ramp
 * e^{i2pi(+-ghost)(x - c)/snx}, an unnormalised length-snx e^{+i2pi} transform (rustfft's inverse;
 * here an iterative radix-2 FFT when snx is a power of two, the direct sum over x otherwise), the bins
 * (kxi - xs) mod snx, and the output centre phase e^{-i2pi m c/snx}.  The FFT's last bits differ from
 * the twiddle sums (as rustfft's do); it is the CPU baseline's `kspace` leg. */
static int g_fft_xstage = 0;
void orc_set_fft_xstage(int on) { g_fft_xstage = on ? 1 : 0; }
typedef struct {
    size_t snx, nx;
    C* tw[2];
    int fft;
    size_t xs;
    C* ramp[2];
    C* centre;
    C* buf;
    C* w;        /* e^{+i2pi k/snx}, k < snx/2 (radix-2) */
} XStage;
static int is_pow2(size_t n) { return n && !(n & (n - 1)); }
static void xstage_new(XStage* xs_, size_t snx, size_t nx, size_t xs, size_t sxs, double xoff, double ghost) {
    double c = (double)sxs + xoff;
    xs_->snx = snx;
    xs_->nx = nx;
    xs_->fft = g_fft_xstage;
    xs_->xs = xs;
    xs_->tw[0] = xs_->tw[1] = NULL;
    xs_->ramp[0] = xs_->ramp[1] = NULL;
    xs_->centre = xs_->buf = xs_->w = NULL;
    if (xs_->fft) {
        for (int p = 0; p < 2; p++) {
            double gh = p == 0 ? ghost : -ghost;
            xs_->ramp[p] = malloc(sizeof(C) * snx);
            for (size_t x = 0; x < snx; x++) xs_->ramp[p][x] = c_cis(TAU * gh * ((double)x - c) / (double)snx);
        }
        xs_->centre = malloc(sizeof(C) * nx);
        for (size_t kxi = 0; kxi < nx; kxi++)
            xs_->centre[kxi] = c_cis(-TAU * ((double)kxi - (double)xs) * c / (double)snx);
        xs_->buf = malloc(sizeof(C) * snx);
        xs_->w = malloc(sizeof(C) * (snx / 2 + 1));
        for (size_t k = 0; k <= snx / 2; k++) xs_->w[k] = c_cis(TAU * (double)k / (double)snx);
        return;
    }
    for (int p = 0; p < 2; p++) {
        double gh = p == 0 ? ghost : -ghost;
        C* t = malloc(sizeof(C) * snx * nx);
        for (size_t kxi = 0; kxi < nx; kxi++) {
            double kx_norm = ((double)kxi - (double)xs + gh) / (double)snx;
            for (size_t x = 0; x < snx; x++) {
                double ph = TAU * kx_norm * ((double)x - c);
                t[x + snx * kxi] = c_cis(ph);
            }
        }
        xs_->tw[p] = t;
    }
}
static void xstage_free(XStage* x) {
    free(x->tw[0]); free(x->tw[1]);
    if (x->fft) { free(x->ramp[0]); free(x->ramp[1]); free(x->centre); free(x->buf); free(x->w); }
}
/* in-place unnormalised e^{+i2pi kx/n} transform of a[0..n) (radix-2 decimation in time) */
static void fft_inverse_pow2(C* a, size_t n, const C* w) {
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { C t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        size_t step = n / len;
        for (size_t i = 0; i < n; i += len) {
            for (size_t k = 0; k < len / 2; k++) {
                C u = a[i + k], v = c_mul(a[i + k + len / 2], w[k * step]);
                a[i + k] = c_add(u, v);
                a[i + k + len / 2].re = u.re - v.re;
                a[i + k + len / 2].im = u.im - v.im;
            }
        }
    }
}
static void xstage_run(const XStage* xs_, const double* g_re, const double* g_im, double ghost_shift, C* out) {
    int pol = ghost_shift < 0.0 ? 1 : 0;
    if (xs_->fft) {
        const size_t n = xs_->snx;
        C* b = xs_->buf;
        for (size_t x = 0; x < n; x++) {
            C g = {g_re[x], g_im[x]};
            b[x] = c_mul(g, xs_->ramp[pol][x]);
        }
        if (is_pow2(n)) {
            fft_inverse_pow2(b, n, xs_->w);
            for (size_t kxi = 0; kxi < xs_->nx; kxi++) {
                long m = (long)kxi - (long)xs_->xs;
                size_t bin = (size_t)(((m % (long)n) + (long)n) % (long)n);
                out[kxi] = c_mul(b[bin], xs_->centre[kxi]);
            }
        } else {
            for (size_t kxi = 0; kxi < xs_->nx; kxi++) {
                long m = (long)kxi - (long)xs_->xs;
                size_t bin = (size_t)(((m % (long)n) + (long)n) % (long)n);
                C acc = {0.0, 0.0};
                for (size_t x = 0; x < n; x++) acc = c_add(acc, c_mul(b[x], c_cis(TAU * (double)((bin * x) % n) / (double)n)));
                out[kxi] = c_mul(acc, xs_->centre[kxi]);
            }
        }
        return;
    }
    const C* tw = xs_->tw[pol];
    for (size_t kxi = 0; kxi < xs_->nx; kxi++) {
        const C* row = tw + xs_->snx * kxi;
        double ar = 0.0, ai = 0.0;
        for (size_t x = 0; x < xs_->snx; x++) {
            double tr = row[x].re, ti = row[x].im;
            ar += g_re[x] * tr - g_im[x] * ti;
            ai += g_re[x] * ti + g_im[x] * tr;
        }
        out[kxi].re = ar;
        out[kxi].im = ai;
    }
}

/* spikes and noise on acquired samples (:997-1039) */
static void spikes_noise(C* k, size_t nx, size_t ny, const mrs_acquisition* a, const uint8_t* mask, uint64_t seed,
                         size_t coil) {
    if (a->n_spikes > 0) {
        C peak = {0.0, 0.0};
        double peak_mag = 0.0;
        for (size_t i = 0; i < nx * ny; i++) {
            double m = c_abs(k[i]);
            if (m > peak_mag) { peak_mag = m; peak = k[i]; }
        }
        C spike = c_scale(peak, a->spike_amplitude);
        size_t nacq = 0;
        for (size_t i = 0; i < nx * ny; i++) nacq += mask[i] != 0;
        if (nacq > 0) {
            size_t* acquired = malloc(sizeof(size_t) * nacq);
            size_t j = 0;
            for (size_t i = 0; i < nx * ny; i++) if (mask[i]) acquired[j++] = i;
            uint64_t rng = seed ^ 0xA5A5123456789ABCULL;
            for (uint64_t s = 0; s < a->n_spikes; s++) {
                size_t pick = acquired[(size_t)orc_rng_next(&rng) % nacq];
                k[pick] = spike;
            }
            free(acquired);
        }
    }
    if (a->noise_variance > 0.0) {
        uint64_t rng = (seed ^ ((uint64_t)coil * 0x9E3779B9ULL)) | 1ULL;
        double sigma = sqrt(a->noise_variance / (double)(nx * ny));
        for (size_t i = 0; i < nx * ny; i++) {
            if (mask[i]) {
                k[i].re += orc_rng_gauss(&rng) * sigma;
                k[i].im += orc_rng_gauss(&rng) * sigma;
            }
        }
    }
}

/* ---------------------------------------------------------------- type-1 NUFFT (src/nufft.rs:25-113)
 * F[k] = sum_j w_j exp(+i 2 pi k v_j / n), k = k_lo .. k_lo + n_out - 1, by gridding with the
 * exponential-of-semicircle kernel at oversampling 2: width w = clamp(ceil(-log10 tol) + 1, 2, 16),
 * beta = 2.30 w, fine grid nf = 2n, spread, inverse DFT of the fine grid (rustfft's unnormalised
 * inverse in the reference; evaluated here directly at the n_out needed bins), divided by the
 * kernel's own transform corr[j] = 1 / (psi(0) + sum_{1 <= m < w/2} 2 psi(m / (w/2)) cos(2 pi k m / nf)).
 * pos: nsrc cells; wre/wim: nw vectors of nsrc; out_re/out_im: nw vectors of n_out. */
static double es_psi(double z, double beta) { return fabs(z) < 1.0 ? exp(beta * (sqrt(1.0 - z * z) - 1.0)) : 0.0; }

int orc_nufft1(size_t n, long k_lo, size_t n_out, double tol, const double* pos, size_t nsrc, const double* wre,
               const double* wim, size_t nw, double* out_re, double* out_im) {
    if (n_out == 0 || k_lo < -(long)(n / 2) || k_lo + (long)n_out - 1 > (long)(n / 2)) return -1;
    long wl = (long)ceil(-log10(tol)) + 1;
    size_t w = (size_t)(wl < 2 ? 2 : (wl > 16 ? 16 : wl));
    const double beta = 2.30 * (double)w;
    const size_t nf = 2 * n;
    const double half = (double)w / 2.0;
    double* corr = malloc(sizeof(double) * n_out);
    for (size_t j = 0; j < n_out; j++) {
        double k = (double)(k_lo + (long)j);
        double sum = es_psi(0.0, beta);
        for (double m = 1.0; m < half; m += 1.0) sum += 2.0 * es_psi(m / half, beta) * cos(TAU * k * m / (double)nf);
        corr[j] = 1.0 / sum;
    }
    double* g = calloc(2 * nf * nw, sizeof(double));
    double* kern = malloc(sizeof(double) * w);
    for (size_t j = 0; j < nsrc; j++) {
        double p = fmod(2.0 * pos[j], (double)nf);
        if (p < 0.0) p += (double)nf;
        long m0 = (long)floor(p - half) + 1;
        for (size_t i = 0; i < w; i++) kern[i] = es_psi(((double)m0 + (double)i - p) / half, beta);
        for (size_t q = 0; q < nw; q++) {
            double wr = wre[q * nsrc + j], wi = wim[q * nsrc + j];
            double* gq = g + 2 * nf * q;
            for (size_t i = 0; i < w; i++) {
                long m = (m0 + (long)i) % (long)nf;
                if (m < 0) m += (long)nf;
                gq[2 * m] += wr * kern[i];
                gq[2 * m + 1] += wi * kern[i];
            }
        }
    }
    for (size_t q = 0; q < nw; q++) {
        const double* gq = g + 2 * nf * q;
        for (size_t j = 0; j < n_out; j++) {
            long bin = (k_lo + (long)j) % (long)nf;
            if (bin < 0) bin += (long)nf;
            double ar = 0.0, ai = 0.0;
            for (size_t m = 0; m < nf; m++) {
                double ph = TAU * (double)(((size_t)bin * m) % nf) / (double)nf;
                double c = cos(ph), sn = sin(ph);
                ar += gq[2 * m] * c - gq[2 * m + 1] * sn;
                ai += gq[2 * m] * sn + gq[2 * m + 1] * c;
            }
            out_re[q * n_out + j] = ar * corr[j];
            out_im[q * n_out + j] = ai * corr[j];
        }
    }
    free(corr); free(g); free(kern);
    return 0;
}

/* The literal sum the NUFFT approximates (the `direct` of src/nufft.rs:301-315). */
void orc_nufft_direct(size_t n, long k_lo, size_t n_out, const double* pos, size_t nsrc, const double* re,
                      const double* im, double* out_re, double* out_im) {
    for (size_t j = 0; j < n_out; j++) {
        double k = (double)(k_lo + (long)j);
        double ar = 0.0, ai = 0.0;
        for (size_t s = 0; s < nsrc; s++) {
            double ph = TAU * k * pos[s] / (double)n;
            double c = cos(ph), sn = sin(ph);
            ar += re[s] * c - im[s] * sn;
            ai += re[s] * sn + im[s] * c;
        }
        out_re[j] = ar;
        out_im[j] = ai;
    }
}

/* ---------------------------------------------------------------- build_coil_kspace_timed (:607-1042),
 * default (std-only) build: rotor y-phase with REANCHOR = 32, per-voxel or per-line relaxation,
 * separable eddy factors, twiddle-table x-stage.  Returns 0, or -1 with the reference's assert. */
typedef struct { size_t dky, par; double dt, dkyn; double* re; double* im; } Rotor;

static int build_timed(const mrs_slice_input* inp, const mrs_acquisition* acq, size_t coil, size_t ncoils,
                       const mrs_line_timing* timing, const double* log_amp, int use_nufft, int* used_nufft,
                       C* kspace) {
    const size_t snx = inp->sim[0], sny = inp->sim[1];
    const size_t nx = inp->acq_matrix[0], ny = inp->acq_matrix[1];
    if (snx % nx != 0 || sny % ny != 0) return -1;
    if (timing->ny != ny) return -2;
    const size_t z = inp->z, nz = inp->nz;
    const size_t Q = inp->n_compartments;
    const double* t_ms = timing->t_ms;
    const double* trf_ms = timing->trf_ms;
    const double* tread_ms = timing->tread_ms;
    double grad[3] = {0.0, 0.0, 0.0};
    if (inp->has_eddy_drive) { grad[0] = inp->eddy_drive[0]; grad[1] = inp->eddy_drive[1]; grad[2] = inp->eddy_drive[2]; }
    const int do_eddy = acq->eddy_strength != 0.0 && inp->has_eddy_drive;
    const int do_eddy_phase = acq->eddy_phase != 0.0 && inp->has_eddy_drive;
    const size_t xs = nx / 2, ys = ny / 2, zs = nz / 2;
    const size_t ox = snx / nx, oy = sny / ny;
    const size_t sxs = ox * xs, sys = oy * ys;
    const double xoff = ((double)ox - 1.0) / 2.0, yoff = ((double)oy - 1.0) / 2.0;
    const size_t nvox = snx * sny;
    uint8_t* mask = malloc(nx * ny);
    orc_sampling_mask(nx, ny, acq, mask);

    enum { REANCHOR = 32 };
    const double zc = (double)z - (double)zs;
    const double zc_c = (double)z - ((double)nz - 1.0) / 2.0;
    double* amp = malloc(sizeof(double) * nvox);
    double* phi0 = malloc(sizeof(double) * nvox);
    double* rate = malloc(sizeof(double) * nvox);
    for (size_t y = 0; y < sny; y++) {
        for (size_t x = 0; x < snx; x++) {
            size_t i = x + snx * y;
            double xa = ((double)x - xoff) / (double)ox, ya = ((double)y - yoff) / (double)oy;
            amp[i] = acq->signal_scale * orc_coil_sensitivity(coil, ncoils, xa, ya, nx, ny);
            double xc = ((double)x - (double)sxs - xoff) / (double)ox;
            double yc = ((double)y - (double)sys - yoff) / (double)oy;
            double p0 = inp->phase0 ? inp->phase0[i] : 0.0;
            if (do_eddy_phase) p0 += acq->eddy_phase * (grad[0] * xc + grad[1] * yc + grad[2] * zc_c);
            if (acq->echo == MRS_ECHO_GRADIENT && acq->do_distortions)
                p0 += TAU * (double)inp->fmap[i] * (acq->t_echo / 1000.0);
            phi0[i] = p0;
            rate[i] = acq->do_distortions ? (double)inp->fmap[i] : 0.0;
        }
    }
#define XC_OF(x) (((double)(x) - (double)sxs - xoff) / (double)ox)
#define YC_OF(y) (((double)(y) - (double)sys - yoff) / (double)oy)
#define KYN_OF(k) (((double)(k) - (double)ys) / (double)sny)

    double* st_re = calloc(nvox, sizeof(double));
    double* st_im = calloc(nvox, sizeof(double));
    Rotor rotors[64];
    size_t nrot = 0;

    int uniform = 1;
    for (size_t c = 0; c < Q; c++) if (inp->t2[c].map) uniform = 0;
    if (inp->t_inhom)
        for (size_t c = 0; c < Q; c++) if (inp->t_inhom[c].map) uniform = 0;
    if (log_amp) uniform = 0;
    double rel[MRS_MAX_COMP + 4];
    for (size_t c = 0; c < Q; c++) rel[c] = 1.0;
    double* ex_re = malloc(sizeof(double) * snx);
    double* ex_im = malloc(sizeof(double) * snx);
    double* ey_re = malloc(sizeof(double) * sny);
    double* ey_im = malloc(sizeof(double) * sny);
    for (size_t x = 0; x < snx; x++) { ex_re[x] = 1.0; ex_im[x] = 0.0; }
    for (size_t y = 0; y < sny; y++) { ey_re[y] = 1.0; ey_im[y] = 0.0; }
    double* g_re = malloc(sizeof(double) * snx);
    double* g_im = malloc(sizeof(double) * snx);
    XStage xst;
    xstage_new(&xst, snx, nx, xs, sxs, xoff, acq->ghost_offset);

    /* NUFFT y-stage (kspace.rs:778-831, feature `kspace`): with t affine in the line index per parity
     * (and the readout polarity the parity's), the per-line y-sum is one type-1 NUFFT per column of
     * the sources displaced to v = (y - c) + sny tau rate, per (compartment, parity) weight vector. */
    double* rows = NULL;   /* [c][re/im][x + snx*kyi] */
    if (use_nufft && !do_eddy && ny >= 3 && uniform) {
        const double tau = (t_ms[2] - t_ms[0]) / 2.0, t0 = t_ms[0];
        const double delta = t_ms[1] - (tau + t0);
        double tmax = 0.0;
        for (size_t k = 0; k < ny; k++) tmax = fmax(tmax, fabs(t_ms[k]));
        tmax = fmax(tmax, 1e-300);
        int affine = 1;
        for (size_t k = 0; k < ny && affine; k++) {
            double pred = tau * (double)k + t0 + ((k % 2 == 1) ? delta : 0.0);
            if (!(fabs(t_ms[k] - pred) <= 1e-9 * tmax) || timing->polarity[k] != ((k % 2 == 1) ? -1 : 1)) affine = 0;
        }
        if (affine) {
            const double tau_s = tau / 1000.0, delta_s = delta / 1000.0;
            const double t_centre = (tau * (double)ys + t0) / 1000.0;
            rows = calloc(2 * Q * snx * ny, sizeof(double));
            double* pos = malloc(sizeof(double) * sny);
            double* wre = malloc(sizeof(double) * 2 * Q * sny);
            double* wim = malloc(sizeof(double) * 2 * Q * sny);
            double* ore = malloc(sizeof(double) * 2 * Q * ny);
            double* oim = malloc(sizeof(double) * 2 * Q * ny);
            for (size_t x = 0; x < snx; x++) {
                for (size_t y = 0; y < sny; y++) {
                    size_t i = x + snx * y;
                    pos[y] = ((double)y - (double)sys - yoff) + (double)sny * tau_s * rate[i];
                    double ph = phi0[i] + TAU * rate[i] * t_centre;
                    double c0 = cos(ph), s0 = sin(ph);
                    double phd = TAU * rate[i] * delta_s;
                    double cd = cos(phd), sd = sin(phd);
                    double c1 = c0 * cd - s0 * sd, s1 = c0 * sd + s0 * cd;
                    for (size_t c = 0; c < Q; c++) {
                        double a = amp[i] * (double)inp->compartments[c][i];
                        wre[(2 * c) * sny + y] = a * c0;
                        wim[(2 * c) * sny + y] = a * s0;
                        wre[(2 * c + 1) * sny + y] = a * c1;
                        wim[(2 * c + 1) * sny + y] = a * s1;
                    }
                }
                orc_nufft1(sny, -(long)ys, ny, 1e-13, pos, sny, wre, wim, 2 * Q, ore, oim);
                for (size_t c = 0; c < Q; c++)
                    for (size_t kyi = 0; kyi < ny; kyi++) {
                        size_t q = 2 * c + (kyi % 2);
                        rows[(2 * c) * snx * ny + x + snx * kyi] = ore[q * ny + kyi];
                        rows[(2 * c + 1) * snx * ny + x + snx * kyi] = oim[q * ny + kyi];
                    }
            }
            free(pos); free(wre); free(wim); free(ore); free(oim);
        }
    }
    if (used_nufft) *used_nufft = rows != NULL;

    for (size_t i = 0; i < nx * ny; i++) { kspace[i].re = 0.0; kspace[i].im = 0.0; }
    int have_prev = 0;
    size_t prev = 0, since_anchor = 0;
    for (size_t kyi = 0; kyi < ny; kyi++) {
        if (!mask[nx * kyi]) continue;
        const double t = t_ms[kyi] / 1000.0;
        const double trf = trf_ms[kyi];
        int anchor = rows == NULL;
        if (rows == NULL && have_prev && since_anchor < REANCHOR) {
            size_t p = prev;
            size_t dky = kyi - p, par = p % 2;
            double dt = t - t_ms[p] / 1000.0, dkyn = KYN_OF(kyi) - KYN_OF(p);
            size_t idx = nrot;
            for (size_t r = 0; r < nrot; r++)
                if (rotors[r].dky == dky && rotors[r].par == par) { idx = r; break; }
            if (idx == nrot) {
                Rotor* r = &rotors[nrot++];
                r->dky = dky; r->par = par; r->dt = dt; r->dkyn = dkyn;
                r->re = malloc(sizeof(double) * nvox);
                r->im = malloc(sizeof(double) * nvox);
                for (size_t y = 0; y < sny; y++) {
                    double yy = (double)y - (double)sys - yoff;
                    for (size_t x = 0; x < snx; x++) {
                        size_t i = x + snx * y;
                        double ph = TAU * (rate[i] * dt + dkyn * yy);
                        r->re[i] = cos(ph);
                        r->im[i] = sin(ph);
                    }
                }
            }
            const Rotor* r = &rotors[idx];
            double adt = fabs(dt) > 1e-300 ? fabs(dt) : 1e-300;
            if (fabs(r->dt - dt) <= 1e-12 * adt && r->dkyn == dkyn) {
                for (size_t i = 0; i < nvox; i++) {
                    double a = st_re[i], b = st_im[i];
                    st_re[i] = a * r->re[i] - b * r->im[i];
                    st_im[i] = a * r->im[i] + b * r->re[i];
                }
                since_anchor += 1;
                anchor = 0;
            }
        }
        if (anchor) {
            /* set_state (:733-745) */
            double kyn = KYN_OF(kyi);
            for (size_t y = 0; y < sny; y++) {
                double yy = (double)y - (double)sys - yoff;
                for (size_t x = 0; x < snx; x++) {
                    size_t i = x + snx * y;
                    double ph = phi0[i] + TAU * (rate[i] * t + kyn * yy);
                    st_re[i] = amp[i] * cos(ph);
                    st_im[i] = amp[i] * sin(ph);
                }
            }
            since_anchor = 0;
        }
        prev = kyi;
        have_prev = 1;

        if (uniform) {
            for (size_t c = 0; c < Q; c++) {
                if (acq->do_relaxation) {
                    double t2c = (double)inp->t2[c].value;
                    double tic = inp->t_inhom ? (double)inp->t_inhom[c].value : acq->t_inhom;
                    if (acq->echo == MRS_ECHO_SPIN) rel[c] = exp(-trf / t2c - fabs(t) * 1000.0 / tic);
                    else rel[c] = exp(-trf / t2c - trf / tic);
                } else {
                    rel[c] = 1.0;
                }
            }
        }
        double ez_re = 1.0, ez_im = 0.0;
        if (do_eddy) {
            double eddy_decay = exp(-tread_ms[kyi] / acq->eddy_tau) * t;
            for (size_t x = 0; x < snx; x++) {
                double xc = XC_OF(x);
                double ph = TAU * (acq->eddy_strength * grad[0] * xc + acq->eddy_quad * grad[0] * xc * xc) * eddy_decay;
                ex_re[x] = cos(ph);
                ex_im[x] = sin(ph);
            }
            for (size_t y = 0; y < sny; y++) {
                double yc = YC_OF(y);
                double ph = TAU * (acq->eddy_strength * grad[1] * yc + acq->eddy_quad * grad[1] * yc * yc) * eddy_decay;
                ey_re[y] = cos(ph);
                ey_im[y] = sin(ph);
            }
            double ph = TAU * (acq->eddy_strength * grad[2] * zc + acq->eddy_quad * grad[2] * zc * zc) * eddy_decay;
            ez_re = cos(ph);
            ez_im = sin(ph);
        }
        for (size_t x = 0; x < snx; x++) { g_re[x] = 0.0; g_im[x] = 0.0; }
        if (rows) {
            for (size_t c = 0; c < Q; c++)
                for (size_t x = 0; x < snx; x++) {
                    g_re[x] += rel[c] * rows[(2 * c) * snx * ny + x + snx * kyi];
                    g_im[x] += rel[c] * rows[(2 * c + 1) * snx * ny + x + snx * kyi];
                }
        }
        for (size_t y = 0; y < sny && !rows; y++) {
            size_t row = snx * y;
            double eyr = ey_re[y], eyi = ey_im[y];
            for (size_t x = 0; x < snx; x++) {
                size_t i = row + x;
                double w = 0.0;
                for (size_t c = 0; c < Q; c++) {
                    double r;
                    if (uniform) {
                        r = rel[c];
                    } else if (acq->do_relaxation) {
                        double t2c = t2_at(&inp->t2[c], i);
                        double tic = inp->t_inhom ? t2_at(&inp->t_inhom[c], i) : acq->t_inhom;
                        if (acq->echo == MRS_ECHO_SPIN) {
                            if (log_amp) r = exp(log_amp[i] - trf / t2c - fabs(t) * 1000.0 / tic);
                            else r = exp(-trf / t2c - fabs(t) * 1000.0 / tic);
                        } else {
                            if (log_amp) r = exp(log_amp[i] - trf / t2c - trf / tic);
                            else r = exp(-trf / t2c - trf / tic);
                        }
                    } else {
                        r = 1.0;
                    }
                    w += r * (double)inp->compartments[c][i];
                }
                double sr = st_re[i], si = st_im[i];
                double mr, mi;
                if (do_eddy) { mr = sr * eyr - si * eyi; mi = sr * eyi + si * eyr; }
                else { mr = sr; mi = si; }
                g_re[x] += w * mr;
                g_im[x] += w * mi;
            }
        }
        if (do_eddy) {
            for (size_t x = 0; x < snx; x++) {
                double a = g_re[x], b = g_im[x];
                double fr = ex_re[x] * ez_re - ex_im[x] * ez_im, fi = ex_re[x] * ez_im + ex_im[x] * ez_re;
                g_re[x] = a * fr - b * fi;
                g_im[x] = a * fi + b * fr;
            }
        }
        double ghost_shift = timing->polarity[kyi] < 0 ? -acq->ghost_offset : acq->ghost_offset;
        xstage_run(&xst, g_re, g_im, ghost_shift, kspace + nx * kyi);
    }
    double n_inv = 1.0 / (double)nvox;
    for (size_t i = 0; i < nx * ny; i++) { kspace[i].re *= n_inv; kspace[i].im *= n_inv; }
    spikes_noise(kspace, nx, ny, acq, mask, inp->slice_seed, coil);

    xstage_free(&xst);
    for (size_t r = 0; r < nrot; r++) { free(rotors[r].re); free(rotors[r].im); }
    free(mask); free(amp); free(phi0); free(rate); free(st_re); free(st_im);
    free(ex_re); free(ex_im); free(ey_re); free(ey_im); free(g_re); free(g_im); free(rows);
    return 0;
#undef XC_OF
#undef YC_OF
#undef KYN_OF
}

/* Which build the per-slice entry points (build_coil_kspace and everything above it: simulate_slice,
 * the GRAPPA reference scan, the series) restate: 0 the default (std-only) build, 1 the `kspace`
 * feature build (NUFFT y-stage where its conditions hold).  Test switch; not thread-safe. */
static int g_kspace_feature = 0;
void orc_set_kspace_feature(int on) { g_kspace_feature = on ? 1 : 0; }

/* build_coil_kspace_timed of the default (std-only) build: the rotor y-stage. */
int orc_build_coil_kspace_timed(const mrs_slice_input* inp, const mrs_acquisition* acq, size_t coil,
                                size_t ncoils, const mrs_line_timing* timing, const double* log_amp, C* kspace) {
    return build_timed(inp, acq, coil, ncoils, timing, log_amp, 0, NULL, kspace);
}

/* The same as the `kspace` feature builds it: the NUFFT y-stage where its conditions hold (uniform
 * relaxation, no eddy polynomial, ny >= 3, affine per-parity timing), else the rotor y-stage;
 * *used_nufft says which ran.  The x-stage stays the twiddle DFT (the feature's FFT x-stage is the
 * same transform). */
int orc_build_coil_kspace_nufft(const mrs_slice_input* inp, const mrs_acquisition* acq, size_t coil, size_t ncoils,
                                int* used_nufft, C* kspace) {
    size_t nx = inp->acq_matrix[0], ny = inp->acq_matrix[1];
    double* t = malloc(sizeof(double) * ny * 3);
    int8_t* pol = malloc(ny);
    orc_line_timing(acq, nx, ny, t, t + ny, t + 2 * ny, pol);
    mrs_line_timing lt = {t, t + ny, t + 2 * ny, pol, ny};
    int rc = build_timed(inp, acq, coil, ncoils, &lt, NULL, 1, used_nufft, kspace);
    free(t); free(pol);
    return rc;
}

/* build_coil_kspace (:597-600): the table from for_acquisition */
int orc_build_coil_kspace(const mrs_slice_input* inp, const mrs_acquisition* acq, size_t coil, size_t ncoils,
                          C* out) {
    size_t nx = inp->acq_matrix[0], ny = inp->acq_matrix[1];
    double* t = malloc(sizeof(double) * ny * 3);
    int8_t* pol = malloc(ny);
    orc_line_timing(acq, nx, ny, t, t + ny, t + 2 * ny, pol);
    mrs_line_timing lt = {t, t + ny, t + 2 * ny, pol, ny};
    int rc = build_timed(inp, acq, coil, ncoils, &lt, NULL, g_kspace_feature, NULL, out);
    free(t);
    free(pol);
    return rc;
}

/* ---------------------------------------------------------------- literal per-line sum (:2225-2406) */
int orc_literal_coil_kspace(const mrs_slice_input* inp, const mrs_acquisition* acq, size_t coil, size_t ncoils,
                            C* kspace) {
    const size_t snx = inp->sim[0], sny = inp->sim[1];
    const size_t nx = inp->acq_matrix[0], ny = inp->acq_matrix[1];
    if (snx % nx != 0 || sny % ny != 0) return -1;
    const size_t z = inp->z, nz = inp->nz, Q = inp->n_compartments;
    Epi e = {nx, ny, acq->t_line, acq->t_echo, acq->reverse_phase};
    double* t_ms = calloc(ny, sizeof(double));
    double* trf_ms = calloc(ny, sizeof(double));
    double* tread_ms = calloc(ny, sizeof(double));
    for (size_t tick = 0; tick < nx * ny; tick++) {
        size_t kx, ky;
        epi_index(&e, tick, &kx, &ky);
        if (kx == nx / 2) {
            t_ms[ky] = epi_t_echo(&e, tick);
            trf_ms[ky] = epi_t_rf(&e, tick);
            tread_ms[ky] = epi_t_prep(&e, tick);
        }
    }
    double skip = (double)pf_skipped(ny, acq) * acq->t_line;
    for (size_t k = 0; k < ny; k++) tread_ms[k] -= skip;
    double grad[3] = {0.0, 0.0, 0.0};
    if (inp->has_eddy_drive) { grad[0] = inp->eddy_drive[0]; grad[1] = inp->eddy_drive[1]; grad[2] = inp->eddy_drive[2]; }
    const int do_eddy = acq->eddy_strength != 0.0 && inp->has_eddy_drive;
    const int do_eddy_phase = acq->eddy_phase != 0.0 && inp->has_eddy_drive;
    const size_t xs = nx / 2, ys = ny / 2, zs = nz / 2;
    const size_t ox = snx / nx, oy = sny / ny;
    const size_t sxs = ox * xs, sys = oy * ys;
    const double xoff = ((double)ox - 1.0) / 2.0, yoff = ((double)oy - 1.0) / 2.0;
    uint8_t* mask = malloc(nx * ny);
    orc_sampling_mask(nx, ny, acq, mask);
    for (size_t i = 0; i < nx * ny; i++) { kspace[i].re = 0.0; kspace[i].im = 0.0; }
    const double n_inv = 1.0 / (double)(snx * sny);
    C* modimg = malloc(sizeof(C) * snx * sny);
    C* g = malloc(sizeof(C) * snx);
    for (size_t kyi = 0; kyi < ny; kyi++) {
        if (!mask[nx * kyi]) continue;
        double t = t_ms[kyi] / 1000.0, trf = trf_ms[kyi];
        double ky_norm = ((double)kyi - (double)ys) / (double)sny;
        double ghost_shift = kyi % 2 == 1 ? -acq->ghost_offset : acq->ghost_offset;
        double eddy_decay = do_eddy ? exp(-tread_ms[kyi] / acq->eddy_tau) * t : 0.0;
        for (size_t y = 0; y < sny; y++) {
            for (size_t x = 0; x < snx; x++) {
                size_t i = x + snx * y;
                double f_real = 0.0;
                for (size_t c = 0; c < Q; c++) {
                    double v = (double)inp->compartments[c][i];
                    if (acq->do_relaxation) {
                        double t2c = t2_at(&inp->t2[c], i);
                        double tic = inp->t_inhom ? t2_at(&inp->t_inhom[c], i) : acq->t_inhom;
                        double inhom_ms = acq->echo == MRS_ECHO_SPIN ? fabs(t) * 1000.0 : trf;
                        v *= exp(-trf / t2c - inhom_ms / tic);
                    }
                    f_real += v;
                }
                double xa = ((double)x - xoff) / (double)ox, ya = ((double)y - yoff) / (double)oy;
                f_real *= acq->signal_scale * orc_coil_sensitivity(coil, ncoils, xa, ya, nx, ny);
                double t_b0 = acq->echo == MRS_ECHO_SPIN ? t : t + acq->t_echo / 1000.0;
                double phi = acq->do_distortions ? (double)inp->fmap[i] * t_b0 : 0.0;
                double phi0 = inp->phase0 ? inp->phase0[i] : 0.0;
                if (do_eddy || do_eddy_phase) {
                    double xc = ((double)x - (double)sxs - xoff) / (double)ox;
                    double yc = ((double)y - (double)sys - yoff) / (double)oy;
                    double zc = (double)z - (double)zs;
                    if (do_eddy) {
                        double lin = grad[0] * xc + grad[1] * yc + grad[2] * zc;
                        double quad = grad[0] * xc * xc + grad[1] * yc * yc + grad[2] * zc * zc;
                        phi += (acq->eddy_strength * lin + acq->eddy_quad * quad) * eddy_decay;
                    }
                    if (do_eddy_phase) {
                        double zc_c = (double)z - ((double)nz - 1.0) / 2.0;
                        phi0 += acq->eddy_phase * (grad[0] * xc + grad[1] * yc + grad[2] * zc_c);
                    }
                }
                modimg[i] = c_scale(c_cis(TAU * phi + phi0), f_real);
            }
        }
        for (size_t x = 0; x < snx; x++) {
            C acc = {0.0, 0.0};
            for (size_t y = 0; y < sny; y++) {
                C ph = c_cis(TAU * ky_norm * ((double)y - (double)sys - yoff));
                acc = c_add(acc, c_mul(modimg[x + snx * y], ph));
            }
            g[x] = acc;
        }
        for (size_t kxi = 0; kxi < nx; kxi++) {
            double kx_norm = ((double)kxi - (double)xs + ghost_shift) / (double)snx;
            C acc = {0.0, 0.0};
            for (size_t x = 0; x < snx; x++)
                acc = c_add(acc, c_mul(g[x], c_cis(TAU * kx_norm * ((double)x - (double)sxs - xoff))));
            kspace[kxi + nx * kyi] = c_scale(acc, n_inv);
        }
    }
    free(modimg); free(g); free(mask); free(t_ms); free(trf_ms); free(tread_ms);
    return 0;
}

/* ---------------------------------------------------------------- reconstruction (:1301-1453, 1457-1633) */

/* KspaceWindow::at (:140-161) */
double orc_window_at(int kind, double a_, double b_, double kx, double ky) {
    double r = 2.0 * sqrt(kx * kx + ky * ky);
    switch (kind) {
    case MRS_WINDOW_NONE: return 1.0;
    case MRS_WINDOW_HANN: return r >= 1.0 ? 0.0 : 0.5 * (1.0 + cos(M_PI * r));
    case MRS_WINDOW_TUKEY: {
        double a = a_ < 0.0 ? 0.0 : (a_ > 1.0 ? 1.0 : a_);
        if (r <= 1.0 - a) return 1.0;
        if (r >= 1.0) return 0.0;
        return 0.5 * (1.0 + cos(M_PI * (r - (1.0 - a)) / (a > 1e-12 ? a : 1e-12)));
    }
    case MRS_WINDOW_FERMI: return 1.0 / (1.0 + exp((r - a_) / (b_ > 1e-12 ? b_ : 1e-12)));
    }
    return 1.0;
}

/* inverse_2d (:1428-1453) */
void orc_inverse_2d(const C* k, size_t nx, size_t ny, C* out) {
    size_t xs = nx / 2, ys = ny / 2;
    C* h = malloc(sizeof(C) * nx * ny);
    for (size_t kxi = 0; kxi < nx; kxi++) {
        for (size_t y = 0; y < ny; y++) {
            C acc = {0.0, 0.0};
            for (size_t kyi = 0; kyi < ny; kyi++) {
                double kyn = ((double)kyi - (double)ys) / (double)ny;
                acc = c_add(acc, c_mul(k[kxi + nx * kyi], c_cis(-TAU * kyn * ((double)y - (double)ys))));
            }
            h[kxi + nx * y] = acc;
        }
    }
    for (size_t y = 0; y < ny; y++) {
        for (size_t x = 0; x < nx; x++) {
            C acc = {0.0, 0.0};
            for (size_t kxi = 0; kxi < nx; kxi++) {
                double kxn = ((double)kxi - (double)xs) / (double)nx;
                acc = c_add(acc, c_mul(h[kxi + nx * y], c_cis(-TAU * kxn * ((double)x - (double)xs))));
            }
            out[x + nx * y] = acc;
        }
    }
    free(h);
}

/* solve_hermitian (:1457-1491); h is n*n row-major, both overwritten */
static void solve_hermitian(C* h, C* b, size_t n, C* x) {
    for (size_t col = 0; col < n; col++) {
        size_t piv = col;
        double best = c_abs(h[col * n + col]);
        for (size_t r = col + 1; r < n; r++) {
            double m = c_abs(h[r * n + col]);
            if (m > best) { best = m; piv = r; }
        }
        if (piv != col) {
            for (size_t c = 0; c < n; c++) { C tmp = h[col * n + c]; h[col * n + c] = h[piv * n + c]; h[piv * n + c] = tmp; }
            C tb = b[col]; b[col] = b[piv]; b[piv] = tb;
        }
        C d = h[col * n + col];
        double m2 = d.re * d.re + d.im * d.im;
        C dinv = m2 < 1e-18 ? (C){0.0, 0.0} : (C){d.re / m2, -d.im / m2};
        for (size_t r = col + 1; r < n; r++) {
            C f = c_mul(h[r * n + col], dinv);
            for (size_t c = col; c < n; c++) {
                C s = c_mul(f, h[col * n + c]);
                h[r * n + c] = c_add(h[r * n + c], c_scale(s, -1.0));
            }
            C sb = c_mul(f, b[col]);
            b[r] = c_add(b[r], c_scale(sb, -1.0));
        }
    }
    for (size_t col = n; col-- > 0;) {
        C s = b[col];
        for (size_t c = col + 1; c < n; c++) s = c_add(s, c_scale(c_mul(h[col * n + c], x[c]), -1.0));
        C d = h[col * n + col];
        double m2 = d.re * d.re + d.im * d.im;
        C dinv = m2 < 1e-18 ? (C){0.0, 0.0} : (C){d.re / m2, -d.im / m2};
        x[col] = c_mul(s, dinv);
    }
}

static const long GRAPPA_DKX[3] = {-1, 0, 1};
#define GRAPPA_LAMBDA 1e-3

/* grappa_feature (:1526-1538); coil[ci] points at coil ci's nx*ny k-space */
static void grappa_feature(C* const* coil, size_t nc, size_t nx, size_t s0, size_t s1, size_t kx, C* f) {
    size_t lines[2] = {s0, s1};
    for (size_t li = 0; li < 2; li++)
        for (size_t ki = 0; ki < 3; ki++) {
            long v = ((long)kx + GRAPPA_DKX[ki]) % (long)nx;
            if (v < 0) v += (long)nx;
            for (size_t ci = 0; ci < nc; ci++) f[(li * 3 + ki) * nc + ci] = coil[ci][(size_t)v + nx * lines[li]];
        }
}

/* GrappaKernel::fit (:1547-1599).  weights: (accel-1) x nc x nfeat, nfeat = 6*nc. */
void orc_grappa_fit(C* const* coil, size_t nc, size_t nx, size_t ny, size_t accel, size_t acs_lines,
                    const uint8_t* lines, C* weights) {
    size_t ys = ny / 2, nfeat = 6 * nc;
    long acs_half = (long)(acs_lines / 2);
    long lo_ = (long)ys - acs_half;
    size_t acs_lo = lo_ > 0 ? (size_t)lo_ : 0;
    size_t acs_hi = (size_t)((long)ys + acs_half);
    if (acs_hi > ny - 1) acs_hi = ny - 1;
    C* aha = malloc(sizeof(C) * nfeat * nfeat);
    C* ahb = malloc(sizeof(C) * nc * nfeat);
    C* f = malloc(sizeof(C) * nfeat);
    C* hcopy = malloc(sizeof(C) * nfeat * nfeat);
    C* bcopy = malloc(sizeof(C) * nfeat);
    for (size_t d = 1; d < accel; d++) {
        C* wd = weights + (d - 1) * nc * nfeat;
        memset(aha, 0, sizeof(C) * nfeat * nfeat);
        memset(ahb, 0, sizeof(C) * nc * nfeat);
        for (size_t b = acs_lo; b < acs_hi; b++) {
            size_t s0 = b, s1 = b + accel, tgt = b + d;
            if (s1 > acs_hi || tgt > acs_hi || !(lines[s0] && lines[s1] && lines[tgt])) continue;
            for (size_t kx = 0; kx < nx; kx++) {
                grappa_feature(coil, nc, nx, s0, s1, kx, f);
                for (size_t i = 0; i < nfeat; i++) {
                    C fic = {f[i].re, -f[i].im};
                    for (size_t j = 0; j < nfeat; j++) aha[i * nfeat + j] = c_add(aha[i * nfeat + j], c_mul(fic, f[j]));
                    for (size_t co = 0; co < nc; co++)
                        ahb[co * nfeat + i] = c_add(ahb[co * nfeat + i], c_mul(fic, coil[co][kx + nx * tgt]));
                }
            }
        }
        double md = -0.0;
        for (size_t i = 0; i < nfeat; i++) md += aha[i * nfeat + i].re;
        double mean_diag = md / (double)nfeat;
        if (mean_diag <= 0.0) {
            memset(wd, 0, sizeof(C) * nc * nfeat);
            continue;
        }
        double s = 1.0 / mean_diag;
        for (size_t i = 0; i < nfeat; i++) {
            for (size_t j = 0; j < nfeat; j++) aha[i * nfeat + j] = c_scale(aha[i * nfeat + j], s);
            aha[i * nfeat + i].re += GRAPPA_LAMBDA;
        }
        for (size_t i = 0; i < nc * nfeat; i++) ahb[i] = c_scale(ahb[i], s);
        for (size_t co = 0; co < nc; co++) {
            memcpy(hcopy, aha, sizeof(C) * nfeat * nfeat);
            memcpy(bcopy, ahb + co * nfeat, sizeof(C) * nfeat);
            solve_hermitian(hcopy, bcopy, nfeat, wd + co * nfeat);
        }
    }
    free(aha); free(ahb); free(f); free(hcopy); free(bcopy);
}

/* GrappaKernel::apply (:1605-1632) */
void orc_grappa_apply(const C* weights, size_t accel, C* const* coil, size_t nc, size_t nx, size_t ny,
                      const uint8_t* acquired) {
    size_t ys = ny / 2, nfeat = 6 * nc;
    C* f = malloc(sizeof(C) * nfeat);
    for (size_t d = 1; d < accel; d++) {
        const C* wd = weights + (d - 1) * nc * nfeat;
        for (size_t ky = 0; ky < ny; ky++) {
            long m = ((long)ky - (long)ys) % (long)accel;
            if (m < 0) m += (long)accel;
            if (m != (long)d || ky < d || ky + accel - d >= ny) continue;
            size_t s0 = ky - d, s1 = ky - d + accel;
            if (!(acquired[s0] && acquired[s1])) continue;
            for (size_t kx = 0; kx < nx; kx++) {
                grappa_feature(coil, nc, nx, s0, s1, kx, f);
                for (size_t co = 0; co < nc; co++) {
                    C v = {0.0, 0.0};
                    for (size_t i = 0; i < nfeat; i++) v = c_add(v, c_mul(wd[co * nfeat + i], f[i]));
                    coil[co][kx + nx * ky] = v;
                }
            }
        }
    }
    free(f);
}

/* reconstruct_coils_full (:1301-1378): GRAPPA (weights may be NULL when accel <= 1), window,
 * inverse_2d, Roemer combine.  coil_kspace (nc*nx*ny) is modified in place.  out: nx*ny. */
void orc_reconstruct(C* coil_kspace, size_t nc, size_t nx, size_t ny, const mrs_acquisition* acq,
                     const C* weights, C* out) {
    size_t xs = nx / 2, ys = ny / 2;
    size_t accel = acq->accel > 1 ? acq->accel : 1;
    C** coil = malloc(sizeof(C*) * nc);
    for (size_t c = 0; c < nc; c++) coil[c] = coil_kspace + c * nx * ny;
    if (accel > 1 && weights) {
        uint8_t* lines = malloc(ny);
        orc_sampling_mask(1, ny, acq, lines);
        orc_grappa_apply(weights, accel, coil, nc, nx, ny, lines);
        free(lines);
    }
    if (acq->window != MRS_WINDOW_NONE) {
        for (size_t c = 0; c < nc; c++)
            for (size_t kyi = 0; kyi < ny; kyi++)
                for (size_t kxi = 0; kxi < nx; kxi++) {
                    double kx = ((double)kxi - (double)xs) / (double)nx;
                    double ky = ((double)kyi - (double)ys) / (double)ny;
                    double w = orc_window_at(acq->window, acq->window_a, acq->window_b, kx, ky);
                    C* k = &coil[c][kxi + nx * kyi];
                    k->re *= w;
                    k->im *= w;
                }
    }
    C* wsum = calloc(nx * ny, sizeof(C));
    double* ssum = calloc(nx * ny, sizeof(double));
    C* img = malloc(sizeof(C) * nx * ny);
    for (size_t c = 0; c < nc; c++) {
        orc_inverse_2d(coil[c], nx, ny, img);
        for (size_t y = 0; y < ny; y++)
            for (size_t x = 0; x < nx; x++) {
                size_t i = x + nx * y;
                double s = orc_coil_sensitivity(c, nc, (double)(i % nx), (double)(i / nx), nx, ny);
                wsum[i] = c_add(wsum[i], c_scale(img[i], s));
                ssum[i] += s * s;
            }
    }
    for (size_t i = 0; i < nx * ny; i++) {
        double s = ssum[i] > 1e-12 ? ssum[i] : 1e-12;
        out[i].re = wsum[i].re / s;
        out[i].im = wsum[i].im / s;
    }
    free(coil); free(wsum); free(ssum); free(img);
}

#define SEED_SALT_GRAPPA_CAL 0x47525043414CULL

/* grappa_reference_kernel (:1256-1262): the slice without eddy or prep drive, read without
 * acceleration or spikes, fitted on its acquired lines.  The caller sets inp->slice_seed.
 * Returns (accel-1) * nc * 6nc weights (malloc), or NULL on a refused input. */
static C* reference_kernel(const mrs_slice_input* inp, const mrs_acquisition* acq) {
    size_t nx = inp->acq_matrix[0], ny = inp->acq_matrix[1];
    size_t nc = acq->n_coils > 1 ? acq->n_coils : 1;
    mrs_slice_input ref = *inp;
    ref.has_eddy_drive = 0;
    ref.has_prep_drive = 0;
    mrs_acquisition acq_ref = *acq;
    acq_ref.accel = 1;
    acq_ref.n_spikes = 0;
    C* coils = malloc(sizeof(C) * nc * nx * ny);
    int rc = 0;
    for (size_t c = 0; c < nc && rc == 0; c++) rc = orc_build_coil_kspace(&ref, &acq_ref, c, nc, coils + c * nx * ny);
    C* weights = NULL;
    if (rc == 0) {
        uint8_t* lines = malloc(ny);
        orc_sampling_mask(1, ny, &acq_ref, lines);
        C** cp = malloc(sizeof(C*) * nc);
        for (size_t c = 0; c < nc; c++) cp[c] = coils + c * nx * ny;
        weights = malloc(sizeof(C) * (acq->accel - 1) * nc * 6 * nc);
        orc_grappa_fit(cp, nc, nx, ny, acq->accel, acq->acs_lines, lines, weights);
        free(cp); free(lines);
    }
    free(coils);
    return weights;
}

/* simulate_slice_with (:1271-1284) with the kernel given (NULL without acceleration): every coil's
 * k-space, then reconstruct_coils_full.  img: nx*ny complex f64 (SliceRecon::combined). */
static int simulate_slice_with(const mrs_slice_input* inp, const mrs_acquisition* acq, const C* weights, C* img) {
    size_t nx = inp->acq_matrix[0], ny = inp->acq_matrix[1];
    size_t nc = acq->n_coils > 1 ? acq->n_coils : 1;
    C* coils = malloc(sizeof(C) * nc * nx * ny);
    int rc = 0;
    for (size_t c = 0; c < nc && rc == 0; c++) rc = orc_build_coil_kspace(inp, acq, c, nc, coils + c * nx * ny);
    if (rc == 0) orc_reconstruct(coils, nc, nx, ny, acq, weights, img);
    free(coils);
    return rc;
}

/* simulate_slice (:1229-1235) via simulate_slice_with (:1271-1284), kernel = None: under
 * acceleration the kernel is fitted on a reference scan of the slice itself, on the slice seed
 * salted by SEED_SALT_GRAPPA_CAL.  out: nx*ny (re, im) f32 pairs. */
int orc_simulate_slice(const mrs_slice_input* inp, const mrs_acquisition* acq, float* out) {
    size_t nx = inp->acq_matrix[0], ny = inp->acq_matrix[1];
    C* weights = NULL;
    if (acq->accel > 1) {
        mrs_slice_input ref = *inp;
        ref.slice_seed = inp->slice_seed ^ SEED_SALT_GRAPPA_CAL;
        if (!(weights = reference_kernel(&ref, acq))) return -1;
    }
    C* img = malloc(sizeof(C) * nx * ny);
    int rc = simulate_slice_with(inp, acq, weights, img);
    if (rc == 0)
        for (size_t i = 0; i < nx * ny; i++) {
            out[2 * i] = (float)img[i].re;
            out[2 * i + 1] = (float)img[i].im;
        }
    free(img); free(weights);
    return rc;
}

/* ---------------------------------------------------------------- object phase (src/phase.rs) */
/* PrepPhase::shot (phase.rs:108-125): one shot's q_eff, translation and rotation, drawn from
 * Rng(seed ^ volume*0x9E37_79B9_7F4A_7C15 ^ slice_group*0xBF58_476D_1CE4_E5B9).
 * shot: q_eff[3], dx[3], rot[3]. */
void orc_prep_shot(const orc_phase_model* m, double magnitude, const double direction[3], uint64_t volume,
                   uint64_t slice_group, uint64_t seed, double shot[9]) {
    for (int i = 0; i < 9; i++) shot[i] = 0.0;
    double n = sqrt(direction[0] * direction[0] + direction[1] * direction[1] + direction[2] * direction[2]);
    if (magnitude <= 0.0 || n < 1e-12) return;
    double s = m->c_q * sqrt(magnitude) / n;
    for (int i = 0; i < 3; i++) shot[i] = direction[i] * s;
    uint64_t rng = seed ^ (volume * 0x9E3779B97F4A7C15ULL) ^ (slice_group * 0xBF58476D1CE4E5B9ULL);
    for (int i = 0; i < 3; i++) shot[3 + i] = orc_rng_gauss(&rng) * m->sigma_dx;
    for (int i = 0; i < 3; i++) shot[6 + i] = orc_rng_gauss(&rng) * m->sigma_rot;
}

/* phase_slice (kspace.rs:1156-1180) with PhaseModel::at (phase.rs:146-148), BackgroundPhase::at
 * (:52-57) and ShotPhase::at (:76-84), in their operation order.  out: snx*sny rad. */
void orc_phase_slice(const orc_phase_model* m, const double shot[9], size_t snx, size_t sny, size_t o, size_t z,
                     size_t nz, double* out) {
    double sxs = (double)snx / 2.0, sys = (double)sny / 2.0;
    double off = ((double)o - 1.0) / 2.0;
    double zc = (double)z - (double)nz / 2.0;
    const double* c = m->coeffs;
    const double* q = shot;
    const double* dx = shot + 3;
    const double* rot = shot + 6;
    for (size_t y = 0; y < sny; y++)
        for (size_t x = 0; x < snx; x++) {
            double r0 = ((double)x - sxs - off) / (double)o;
            double r1 = ((double)y - sys - off) / (double)o;
            double r2 = zc;
            double bg = c[0] + c[1] * r0 + c[2] * r1 + c[3] * r2 + c[4] * r0 * r0 + c[5] * r1 * r1 + c[6] * r2 * r2 +
                        c[7] * r0 * r1 + c[8] * r0 * r2 + c[9] * r1 * r2;
            double cr0 = rot[1] * r2 - rot[2] * r1;
            double cr1 = rot[2] * r0 - rot[0] * r2;
            double cr2 = rot[0] * r1 - rot[1] * r0;
            double sh = -0.0;
            sh += q[0] * (dx[0] + cr0);
            sh += q[1] * (dx[1] + cr1);
            sh += q[2] * (dx[2] + cr2);
            out[x + snx * y] = m->global + bg + sh;
        }
}

/* ---------------------------------------------------------------- series (acquire_volumes) */
/* acquire_volumes_complex (:1851-2125) with default options (every slice local, no capture, one
 * echo time), as simulate_acquisition_oversampled (:1687-1718) calls it.
 *   images[q]: (x + snx*(y + sny*z))*nv + g; t2[q] / t_inhom[q] (t_inhom may be NULL): a map over
 *   the whole simulation volume (snx*sny*nz) or a uniform value; fmap snx*sny*nz;
 *   eddy[g]: has_eddy[g] ? eddy_drive[3g..3g+2] : None; prep[g]: has_prep[g] ? (prep_mag[g],
 *   prep_dir[3g..]) : None; model: the object phase (required: the reference always passes
 *   phase_slice's values as phase0; zero coefficients give a real object); noise_sigma: nx*ny*nz
 *   or NULL.
 * GRAPPA (:1940-1985): one kernel per slice fitted on volume 0's images with no shot phase and
 * grappa_cal_slice_seed(z, receiver_seed) (:1655-1657).
 * Output re/im: nx*ny*nz*nv f32, (x + nx*(y + ny*z))*nv + g.  Returns 0, or -1 on a refused
 * slice. */
int orc_acquire_volumes(const size_t sim_dims[3], const size_t acq_dims[3], size_t nv, const float* const* images,
                        size_t ncomp, const mrs_t2_slice* t2, const mrs_t2_slice* t_inhom, const float* fmap,
                        const mrs_acquisition* acq, const int32_t* has_eddy, const double* eddy_drive,
                        const int32_t* has_prep, const double* prep_mag, const double* prep_dir,
                        const orc_phase_model* model, uint64_t excitation_seed, uint64_t receiver_seed,
                        const float* noise_sigma, float* re_out, float* im_out) {
    size_t snx = sim_dims[0], sny = sim_dims[1], nz = sim_dims[2];
    size_t nx = acq_dims[0], ny = acq_dims[1];
    if (acq_dims[2] != nz || snx % nx || sny % ny || snx / nx != sny / ny || ncomp > MRS_MAX_COMP) return -1;
    size_t o = snx / nx, P = snx * sny;
    size_t nvox_acq = nx * ny * nz;
    double zero_shot[9] = {0};
    double* phi = malloc(sizeof(double) * P);
    float* cs = malloc(sizeof(float) * ncomp * P);
    const float* refs[MRS_MAX_COMP];
    mrs_t2_slice t2s[MRS_MAX_COMP], tis[MRS_MAX_COMP];
    C* img = malloc(sizeof(C) * nx * ny);
    C** kernels = calloc(nz, sizeof(C*));
    int rc = 0;
    /* slice inputs of local slice zl with plane pointers into `planes` (compartment images) */
#define SLICE_AT(zl, planes, fmap_plane)                                                              \
    mrs_slice_input inp;                                                                                \
    memset(&inp, 0, sizeof inp);                                                                        \
    for (size_t q = 0; q < ncomp; q++) {                                                                \
        refs[q] = (planes) + q * P;                                                                     \
        t2s[q] = t2[q];                                                                                 \
        if (t2[q].map) t2s[q].map = t2[q].map + (zl) * P;                                               \
        if (t_inhom) { tis[q] = t_inhom[q]; if (t_inhom[q].map) tis[q].map = t_inhom[q].map + (zl) * P; } \
    }                                                                                                   \
    inp.compartments = refs; inp.n_compartments = ncomp; inp.t2 = t2s; inp.t_inhom = t_inhom ? tis : NULL; \
    inp.fmap = (fmap_plane); inp.phase0 = phi; inp.sim[0] = snx; inp.sim[1] = sny;                      \
    inp.acq_matrix[0] = nx; inp.acq_matrix[1] = ny; inp.z = (zl); inp.nz = nz;
    if (acq->accel > 1 && nv > 0) {
        for (size_t zl = 0; zl < nz && rc == 0; zl++) {
            for (size_t q = 0; q < ncomp; q++)
                for (size_t i = 0; i < P; i++) cs[q * P + i] = images[q][(zl * P + i) * nv + 0];
            orc_phase_slice(model, zero_shot, snx, sny, o, zl, nz, phi);
            SLICE_AT(zl, cs, fmap + zl * P)
            inp.slice_seed = (zl * 0x9E37ULL) ^ receiver_seed ^ SEED_SALT_GRAPPA_CAL;
            if (!(kernels[zl] = reference_kernel(&inp, acq))) rc = -1;
        }
    }
    for (size_t g = 0; g < nv && rc == 0; g++) {
        for (size_t zl = 0; zl < nz && rc == 0; zl++) {
            for (size_t q = 0; q < ncomp; q++)
                for (size_t i = 0; i < P; i++) cs[q * P + i] = images[q][(zl * P + i) * nv + g];
            double shot[9] = {0};
            if (model->has_prep && has_prep && has_prep[g])
                orc_prep_shot(model, prep_mag[g], prep_dir + 3 * g, g, zl, excitation_seed, shot);
            orc_phase_slice(model, shot, snx, sny, o, zl, nz, phi);
            SLICE_AT(zl, cs, fmap + zl * P)
            if (has_eddy && has_eddy[g]) {
                inp.has_eddy_drive = 1;
                for (int k = 0; k < 3; k++) inp.eddy_drive[k] = eddy_drive[3 * g + k];
            }
            if (has_prep && has_prep[g]) {
                inp.has_prep_drive = 1;
                inp.prep_magnitude = prep_mag[g];
                for (int k = 0; k < 3; k++) inp.prep_direction[k] = prep_dir[3 * g + k];
            }
            inp.slice_seed = (((uint64_t)g * 0x1000001ULL + zl) * 0x9E37ULL) ^ receiver_seed;
            rc = simulate_slice_with(&inp, acq, acq->accel > 1 ? kernels[zl] : NULL, img);
            for (size_t y = 0; y < ny && rc == 0; y++)
                for (size_t x = 0; x < nx; x++) {
                    C c = img[x + nx * y];
                    float re = (float)c.re, im = (float)c.im;
                    size_t vox = x + nx * (y + ny * zl);
                    if (noise_sigma) {
                        double sd = (double)noise_sigma[vox];
                        if (sd > 0.0) {
                            uint64_t rng = (receiver_seed ^ ((uint64_t)g * 0x9E3779B97F4A7C15ULL) ^
                                            ((uint64_t)vox * 0xBF58476D1CE4E5B9ULL)) | 1ULL;
                            re += (float)(orc_rng_gauss(&rng) * sd);
                            im += (float)(orc_rng_gauss(&rng) * sd);
                        }
                    }
                    re_out[vox * nv + g] = re;
                    im_out[vox * nv + g] = im;
                }
        }
    }
#undef SLICE_AT
    for (size_t zl = 0; zl < nz; zl++) free(kernels[zl]);
    free(kernels); free(phi); free(cs); free(img);
    (void)nvox_acq;
    return rc;
}

/* acquire_volumes (:1840-1846): magnitude sqrt(re*re + im*im) and phase atan2(im, re) in f32. */
void orc_mag_phase(const float* re, const float* im, size_t n, float* mag, float* ph) {
    for (size_t i = 0; i < n; i++) {
        float r = re[i], m = im[i];
        mag[i] = sqrtf(r * r + m * m);
        ph[i] = atan2f(m, r);
    }
}

/* The series inputs of tests/golden_baseline.rs:91-116 (16x16x3, o = 2, 4 volumes, two
 * compartments from one Rng(0xBA5E) stream).  images: 2 * nsim*nv in the entry point's layout
 * (vox*nv + g); fmap, t2map: nsim. */
void orc_series_inputs(float* images, float* fmap, float* t2map) {
    const size_t nx = 16, ny = 16, nz = 3, o = 2, nv = 4;
    const size_t snx = nx * o, sny = ny * o, nsim = snx * sny * nz;
    uint64_t rng = 0xBA5E;
    for (size_t c = 0; c < 2; c++)
        for (size_t i = 0; i < nsim * nv; i++) {
            size_t v = i / nsim, vox = i % nsim;
            double x = (double)(vox % snx) / (double)snx, y = (double)((vox / snx) % sny) / (double)sny;
            double dx = x - 0.5, dy = y - 0.5;
            double blob = exp(-((dx * dx + dy * dy) / 0.05));
            double val = (1.0 + 0.1 * (double)c + 0.01 * (double)v) * blob + 0.02 * orc_rng_unit(&rng);
            images[c * nsim * nv + vox * nv + v] = (float)val;
        }
    for (size_t i = 0; i < nsim; i++) {
        fmap[i] = (float)(40.0 * ((double)(i % snx) / (double)snx - 0.5));
        t2map[i] = (float)(50.0 + (double)(i % 7) * 10.0);
    }
}

/* ---------------------------------------------------------------- golden inputs (tests/golden_baseline.rs:32-55) */
/* The three shapes (17x70 o2, 16x40 o1, 8x9 o3) share one Rng(0xC0FFEE) stream, so the inputs of
 * shape `s` are drawn after those of the shapes before it.  comps: 3*n, fmap n, phase0 n,
 * t2_map n, ti_map n; n = snx*sny. */
void orc_golden_slice_inputs(int shape, float* comps, float* fmap, double* phase0, float* t2_map, float* ti_map) {
    static const size_t shapes[3][3] = {{17, 70, 2}, {16, 40, 1}, {8, 9, 3}};
    uint64_t rng = 0xC0FFEE;
    for (int s = 0; s <= shape; s++) {
        size_t nx = shapes[s][0], ny = shapes[s][1], o = shapes[s][2];
        size_t snx = nx * o, sny = ny * o, n = snx * sny;
        int keep = s == shape;
        for (size_t c = 0; c < 3; c++)
            for (size_t i = 0; i < n; i++) {
                double x = (double)(i % snx) / (double)snx, y = (double)(i / snx) / (double)sny;
                double dx = x - 0.5, dy = y - 0.45;
                double blob = exp(-((dx * dx + dy * dy) / (0.03 * (double)(c + 1))));
                float v = (float)(blob + 0.05 * orc_rng_unit(&rng));
                if (keep) comps[c * n + i] = v;
            }
        for (size_t i = 0; i < n; i++) {
            double x = (double)(i % snx) / (double)snx, y = (double)(i / snx) / (double)sny;
            float v = (float)(300.0 * (sin(3.0 * x) * cos(2.0 * y + 1.0)) + 20.0 * orc_rng_unit(&rng));
            if (keep) fmap[i] = v;
        }
        for (size_t i = 0; i < n; i++) {
            double v = TAU * orc_rng_unit(&rng);
            if (keep) phase0[i] = v;
        }
        if (keep) {
            for (size_t i = 0; i < n; i++) {
                t2_map[i] = (float)(60.0 + 80.0 * ((double)(i % snx) / (double)snx));
                ti_map[i] = (float)(30.0 + 40.0 * ((double)(i / snx) / (double)sny));
            }
        }
    }
}

/* box_hires (:497-513) */
void orc_box_hires(size_t snx, size_t sny, double x0, double x1, double y0, double y1, float* v) {
    for (size_t i = 0; i < snx * sny; i++) v[i] = 0.0f;
    for (size_t y = 0; y < sny; y++) {
        double lo = (double)y, hi = (double)y + 1.0;
        double fy = (hi < y1 ? hi : y1) - (lo > y0 ? lo : y0);
        fy = fy < 0.0 ? 0.0 : (fy > 1.0 ? 1.0 : fy);
        if (fy <= 0.0) continue;
        for (size_t x = 0; x < snx; x++) {
            double xl = (double)x, xh = (double)x + 1.0;
            double fx = (xh < x1 ? xh : x1) - (xl > x0 ? xl : x0);
            fx = fx < 0.0 ? 0.0 : (fx > 1.0 ? 1.0 : fx);
            v[x + snx * y] = (float)(fx * fy);
        }
    }
}

/* ---------------------------------------------------------------- direct FP64 forward, plan inputs
 * The forward of build_coil_kspace_timed (src/kspace.rs:607-994) evaluated directly per (voxel,
 * line) in FP64 -- the literal sum's closed-form phase (src/kspace.rs:2345-2383) -- on the inputs of
 * the batched plan API: explicit complex coil maps (mrs_c32, times signal_scale), real or complex
 * compartment images, T2/T2' maps or uniforms, optional timing table and log_amp.  Host pointers in
 * the descriptors.  No spikes or noise.  out: C x ny x nx.  Pinned against the restatement above
 * (ring maps) and by closed forms (tests/test_oracle.py). */
int orc_forward_direct(const mrs_group_desc* g, const mrs_slice_desc* s, const void* images, C* out) {
    const size_t snx = g->snx, sny = g->sny, nx = g->nx, ny = g->ny, Q = g->n_compartments, NC = g->n_coils;
    if (snx % nx != 0 || sny % ny != 0) return -1;
    const mrs_acquisition* acq = &g->acq;
    const size_t P = snx * sny;
    double* t_ms = malloc(sizeof(double) * ny * 3);
    double* trf_ms = t_ms + ny;
    double* tread_ms = t_ms + 2 * ny;
    int8_t* pol = malloc(ny);
    if (g->timing) {
        memcpy(t_ms, g->timing->t_ms, ny * 8);
        memcpy(trf_ms, g->timing->trf_ms, ny * 8);
        memcpy(tread_ms, g->timing->tread_ms, ny * 8);
        memcpy(pol, g->timing->polarity, ny);
    } else {
        orc_line_timing(acq, nx, ny, t_ms, trf_ms, tread_ms, pol);
    }
    uint8_t* mask = malloc(nx * ny);
    orc_sampling_mask(nx, ny, acq, mask);
    const int eddy_on = acq->eddy_strength != 0.0 && s->has_eddy_drive;
    const int eddy_ph = acq->eddy_phase != 0.0 && s->has_eddy_drive;
    const double g0 = s->eddy_drive[0], g1 = s->eddy_drive[1], g2 = s->eddy_drive[2];
    const size_t xs = nx / 2, ys = ny / 2, zs = s->nz / 2;
    const size_t ox = snx / nx, oy = sny / ny;
    const size_t sxs = ox * xs, sys = oy * ys;
    const double xoff = ((double)ox - 1.0) / 2.0, yoff = ((double)oy - 1.0) / 2.0;
    const double zc = (double)s->z - (double)zs, zc_c = (double)s->z - ((double)s->nz - 1.0) / 2.0;
    double* phi0 = malloc(sizeof(double) * P);
    double* poly = malloc(sizeof(double) * P);
    double* rate = malloc(sizeof(double) * P);
    for (size_t y = 0; y < sny; y++)
        for (size_t x = 0; x < snx; x++) {
            size_t i = x + snx * y;
            double xc = ((double)x - (double)sxs - xoff) / (double)ox, yc = ((double)y - (double)sys - yoff) / (double)oy;
            double p0 = s->phase0 ? s->phase0[i] : 0.0;
            if (eddy_ph) p0 += acq->eddy_phase * (g0 * xc + g1 * yc + g2 * zc_c);
            if (acq->echo == MRS_ECHO_GRADIENT && acq->do_distortions) p0 += TAU * (double)s->fmap[i] * (acq->t_echo / 1000.0);
            phi0[i] = p0;
            rate[i] = acq->do_distortions ? (double)s->fmap[i] : 0.0;
            poly[i] = eddy_on ? acq->eddy_strength * (g0 * xc + g1 * yc + g2 * zc) +
                                    acq->eddy_quad * (g0 * xc * xc + g1 * yc * yc + g2 * zc * zc)
                              : 0.0;
        }
    C* G = malloc(sizeof(C) * NC * snx);
    const double cctr = (double)sxs + xoff, ninv = 1.0 / (double)P;
    for (size_t i = 0; i < NC * nx * ny; i++) { out[i].re = 0.0; out[i].im = 0.0; }
    for (size_t ky = 0; ky < ny; ky++) {
        if (!mask[nx * ky]) continue;
        const double t = t_ms[ky] / 1000.0, trf = trf_ms[ky];
        const double tau = acq->echo == MRS_ECHO_SPIN ? fabs(t) * 1000.0 : trf;
        const double d = eddy_on ? exp(-tread_ms[ky] / acq->eddy_tau) * t : 0.0;
        const double kyn = ((double)ky - (double)ys) / (double)sny;
        memset(G, 0, sizeof(C) * NC * snx);
        for (size_t y = 0; y < sny; y++) {
            const double yy = (double)y - (double)sys - yoff;
            for (size_t x = 0; x < snx; x++) {
                const size_t i = x + snx * y;
                C w = {0.0, 0.0};
                for (size_t q = 0; q < Q; q++) {
                    double r = 1.0;
                    if (acq->do_relaxation) {
                        double t2 = s->t2 ? (double)s->t2[q * P + i] : (double)s->t2_uniform[q];
                        double ti = s->t_inhom ? (double)s->t_inhom[q * P + i]
                                               : (s->has_t_inhom ? (double)s->t_inhom_uniform[q] : acq->t_inhom);
                        r = exp(-trf / t2 - tau / ti);
                    }
                    if (g->complex_images) {
                        const float* m = (const float*)images + 2 * (q * P + i);
                        w.re += r * (double)m[0];
                        w.im += r * (double)m[1];
                    } else {
                        w.re += r * (double)((const float*)images)[q * P + i];
                    }
                }
                if (s->log_amp) { double e = exp(s->log_amp[i]); w.re *= e; w.im *= e; }
                const double ph = phi0[i] + TAU * (rate[i] * t + kyn * yy + poly[i] * d);
                const C e = c_mul(w, c_cis(ph));
                for (size_t c = 0; c < NC; c++) {
                    const mrs_c32 sv = s->coil_maps[c * P + i];
                    const C sc = {(double)sv.re * acq->signal_scale, (double)sv.im * acq->signal_scale};
                    G[c * snx + x] = c_add(G[c * snx + x], c_mul(sc, e));
                }
            }
        }
        const double shift = pol[ky] < 0 ? -acq->ghost_offset : acq->ghost_offset;
        const double gh = (shift < 0.0) ? -acq->ghost_offset : acq->ghost_offset;
        for (size_t kx = 0; kx < nx; kx++) {
            const double kxn = ((double)kx - (double)xs + gh) / (double)snx;
            for (size_t c = 0; c < NC; c++) {
                C acc = {0.0, 0.0};
                for (size_t x = 0; x < snx; x++) acc = c_add(acc, c_mul(G[c * snx + x], c_cis(TAU * kxn * ((double)x - cctr))));
                out[(c * ny + ky) * nx + kx] = c_scale(acc, ninv);
            }
        }
    }
    free(t_ms); free(pol); free(mask); free(phi0); free(poly); free(rate); free(G);
    return 0;
}
