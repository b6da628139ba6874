/* sanitize_check.c -- drives the oracle (mrsim_oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer (test infrastructure; built and run by `make -C oracle sanitize`).
 *
 * Every entry point the tests use, on the reference's 13 golden configurations
 * (tests/golden_baseline.rs:56-83) at the three golden shapes: the full simulate_slice pipeline
 * (k-spaces, GRAPPA calibration, reconstruction), the literal per-line sum, the `kspace` feature
 * build (NUFFT y-stage), the FFT x-stage, the timing tables, the partial-Fourier and GRAPPA masks,
 * and the type-1 NUFFT.  A sanitizer finding aborts with a non-zero status; the run prints a checksum
 * so the work is not optimised away. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mrsim_gpu.h"

void orc_acquisition_default(mrs_acquisition* a);
void orc_line_timing(const mrs_acquisition* a, size_t nx, size_t ny, double* t, double* trf, double* tread,
                     int8_t* pol);
void orc_sampling_mask(size_t nx, size_t ny, const mrs_acquisition* a, uint8_t* m);
void orc_golden_slice_inputs(int shape, float* comps, float* fmap, double* phase0, float* t2_map, float* ti_map);
int orc_simulate_slice(const mrs_slice_input* inp, const mrs_acquisition* acq, float* out);
int orc_build_coil_kspace(const mrs_slice_input* inp, const mrs_acquisition* acq, size_t coil, size_t ncoils,
                          mrs_c64* out);
int orc_literal_coil_kspace(const mrs_slice_input* inp, const mrs_acquisition* acq, size_t coil, size_t ncoils,
                            mrs_c64* out);
void orc_set_kspace_feature(int on);
void orc_set_fft_xstage(int on);
int orc_nufft1(size_t n, long k_lo, size_t n_out, double tol, const double* pos, size_t nsrc, const double* wre,
               const double* wim, size_t nw, double* ore, double* oim);

static const size_t SHAPES[3][3] = {{17, 70, 2}, {16, 40, 1}, {8, 9, 3}};

typedef struct {
    const char* name;
    int do_dist, do_relax, reverse, pf_mode, accel, acs, n_coils, maps, has_eddy;
    double ghost, eddy_strength, eddy_quad, eddy_tau, eddy_phase, pf, noise;
    double eddy[3];
} Case;

static const Case CASES[] = {
    {"clean", 0, 0, 0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1.0, 0, {0, 0, 0}},
    {"distortion+relaxation", 1, 1, 0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1.0, 0, {0, 0, 0}},
    {"reverse", 1, 1, 1, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1.0, 0, {0, 0, 0}},
    {"ghost", 1, 1, 0, 0, 1, 0, 1, 0, 0, 0.015, 0, 0, 0, 0, 1.0, 0, {0, 0, 0}},
    {"eddy-poly", 1, 1, 0, 0, 1, 0, 1, 0, 1, 0, 3.0, 0.4, 70.0, 0, 1.0, 0, {0.3, -0.8, 0.5}},
    {"eddy-phase", 1, 1, 0, 0, 1, 0, 1, 0, 1, 0, 0, 0, 0, 0.2, 1.0, 0, {0.6, 0.6, 0.5}},
    {"pf-fiberfox", 1, 1, 0, MRS_PF_FIBERFOX, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0.75, 0, {0, 0, 0}},
    {"pf-contiguous-reverse", 1, 1, 1, MRS_PF_CONTIGUOUS, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0.75, 0, {0, 0, 0}},
    {"grappa-coils", 1, 1, 0, 0, 2, 6, 4, 0, 0, 0, 0, 0, 0, 0, 1.0, 0, {0, 0, 0}},
    {"grappa3", 1, 1, 0, 0, 3, 8, 4, 0, 0, 0, 0, 0, 0, 0, 1.0, 0, {0, 0, 0}},
    {"everything", 1, 1, 0, 0, 2, 8, 3, 0, 1, 0.02, 2.0, 0.3, 0, 0.1, 0.8, 0, {0.5, 0.5, 0.7}},
    {"maps", 1, 1, 0, 0, 1, 0, 1, 1, 0, 0, 0, 0, 0, 0, 1.0, 0, {0, 0, 0}},
    {"maps-coils-noise", 1, 1, 0, 0, 1, 0, 3, 1, 0, 0, 0, 0, 0, 0, 1.0, 4.0, {0, 0, 0}},
};

static double checksum(const void* p, size_t bytes) {
    const unsigned char* b = p;
    double s = 0;
    for (size_t i = 0; i < bytes; i++) s += b[i] * (double)((i % 7) + 1);
    return s;
}

int main(void) {
    double sum = 0;
    for (int shape = 0; shape < 3; shape++) {
        const size_t nx = SHAPES[shape][0], ny = SHAPES[shape][1], o = SHAPES[shape][2];
        const size_t snx = nx * o, sny = ny * o, n = snx * sny;
        float* comps = malloc(sizeof(float) * 3 * n);
        float* fmap = malloc(sizeof(float) * n);
        double* phase0 = malloc(sizeof(double) * n);
        float* t2m = malloc(sizeof(float) * n);
        float* tim = malloc(sizeof(float) * n);
        orc_golden_slice_inputs(shape, comps, fmap, phase0, t2m, tim);
        const float* cp[3] = {comps, comps + n, comps + 2 * n};
        for (size_t k = 0; k < sizeof CASES / sizeof CASES[0]; k++) {
            const Case* c = &CASES[k];
            mrs_acquisition acq;
            orc_acquisition_default(&acq);
            acq.do_distortions = c->do_dist;
            acq.do_relaxation = c->do_relax;
            acq.signal_scale = 100.0;
            acq.reverse_phase = c->reverse;
            acq.pf_mode = c->pf_mode;
            acq.partial_fourier = c->pf;
            acq.accel = (uint64_t)c->accel;
            acq.acs_lines = (uint64_t)c->acs;
            acq.n_coils = (uint64_t)c->n_coils;
            acq.ghost_offset = c->ghost;
            acq.eddy_strength = c->eddy_strength;
            acq.eddy_quad = c->eddy_quad;
            if (c->eddy_tau > 0) acq.eddy_tau = c->eddy_tau;
            acq.eddy_phase = c->eddy_phase;
            acq.noise_variance = c->noise;
            mrs_t2_slice t2[3], ti[3];
            if (c->maps) {
                t2[0] = (mrs_t2_slice){t2m, 0.0f}; t2[1] = (mrs_t2_slice){NULL, 100.0f}; t2[2] = (mrs_t2_slice){t2m, 0.0f};
                ti[0] = (mrs_t2_slice){tim, 0.0f}; ti[1] = (mrs_t2_slice){NULL, 45.0f}; ti[2] = (mrs_t2_slice){tim, 0.0f};
            } else {
                t2[0] = (mrs_t2_slice){NULL, 70.0f}; t2[1] = (mrs_t2_slice){NULL, 100.0f}; t2[2] = (mrs_t2_slice){NULL, 2000.0f};
            }
            mrs_slice_input inp;
            memset(&inp, 0, sizeof inp);
            inp.compartments = cp;
            inp.n_compartments = 3;
            inp.t2 = t2;
            inp.t_inhom = c->maps ? ti : NULL;
            inp.fmap = fmap;
            inp.phase0 = phase0;
            inp.sim[0] = snx; inp.sim[1] = sny;
            inp.acq_matrix[0] = nx; inp.acq_matrix[1] = ny;
            inp.z = 3; inp.nz = 9;
            inp.has_eddy_drive = c->has_eddy;
            for (int i = 0; i < 3; i++) inp.eddy_drive[i] = c->eddy[i];
            inp.slice_seed = 7;
            /* the whole pipeline: coil k-spaces, GRAPPA calibration and apply, reconstruction */
            float* img = malloc(sizeof(float) * 2 * nx * ny);
            if (orc_simulate_slice(&inp, &acq, img) != 0) { fprintf(stderr, "%s refused\n", c->name); return 2; }
            sum += checksum(img, sizeof(float) * 2 * nx * ny);
            free(img);
            /* one coil through each build: default, `kspace` feature, FFT x-stage, literal sum */
            mrs_c64* kbuf = malloc(sizeof(mrs_c64) * nx * ny);
            const size_t nc = c->n_coils > 0 ? (size_t)c->n_coils : 1;
            orc_build_coil_kspace(&inp, &acq, nc - 1, nc, kbuf);
            sum += checksum(kbuf, sizeof(mrs_c64) * nx * ny);
            orc_set_kspace_feature(1);
            orc_build_coil_kspace(&inp, &acq, 0, nc, kbuf);
            orc_set_kspace_feature(0);
            sum += checksum(kbuf, sizeof(mrs_c64) * nx * ny);
            orc_set_fft_xstage(1);
            orc_build_coil_kspace(&inp, &acq, 0, nc, kbuf);
            orc_set_fft_xstage(0);
            sum += checksum(kbuf, sizeof(mrs_c64) * nx * ny);
            if (shape == 2) {
                orc_literal_coil_kspace(&inp, &acq, 0, nc, kbuf);
                sum += checksum(kbuf, sizeof(mrs_c64) * nx * ny);
            }
            free(kbuf);
            /* timing table and mask */
            double* t = malloc(sizeof(double) * 3 * ny);
            int8_t* pol = malloc(ny);
            uint8_t* mask = malloc(nx * ny);
            orc_line_timing(&acq, nx, ny, t, t + ny, t + 2 * ny, pol);
            orc_sampling_mask(nx, ny, &acq, mask);
            sum += checksum(t, sizeof(double) * 3 * ny) + checksum(mask, nx * ny);
            free(t); free(pol); free(mask);
        }
        free(comps); free(fmap); free(phase0); free(t2m); free(tim);
    }
    /* the type-1 NUFFT on random sources, with a band straddling zero */
    {
        const size_t nsrc = 37, nw = 2, n = 48, n_out = 20;
        double pos[37], wre[74], wim[74], ore[40], oim[40];
        uint64_t s = 12345;
        for (size_t i = 0; i < nsrc; i++) { s = s * 6364136223846793005ULL + 1; pos[i] = (double)(s >> 11) / 9007199254740992.0 * n; }
        for (size_t i = 0; i < nw * nsrc; i++) { s = s * 6364136223846793005ULL + 1; wre[i] = (double)(s >> 40) / 16777216.0 - 0.5; wim[i] = 0.25 - wre[i]; }
        if (orc_nufft1(n, -10, n_out, 1e-13, pos, nsrc, wre, wim, nw, ore, oim) != 0) { fprintf(stderr, "nufft refused\n"); return 2; }
        sum += checksum(ore, sizeof ore) + checksum(oim, sizeof oim);
    }
    printf("sanitize_check ok %.6e\n", sum);
    return 0;
}
