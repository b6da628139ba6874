/*
 * mrsim_gpu.h -- C ABI of the MI355X exact k-space DFT forward model (and its adjoint).
 *
 * The reference (PennLINC/mrsim-acq, Rust) has no FFI layer: its seam is the crate-internal
 *   fn build_coil_kspace_timed(inp: &SliceInput, acq: &Acquisition, coil: usize, ncoils: usize,
 *                              timing: &LineTiming, log_amp: Option<&[f64]>) -> Vec<C>
 * (src/kspace.rs:607-609).  This header is the plain-C mirror of that call and of the POD types it
 * reads, plus batched plan/execute entry points that run many (slice, volume, echo) units in one
 * launch on one GPU.  Everything here is plain C: pointers and sizes, no C++ or torch types.
 *
 * Conventions (identical to the reference):
 *   images on the simulation grid, index x + snx*y;  k-space on the acquired matrix, kx + nx*ky;
 *   forward sign e^{+i2pi(...)} (src/kspace.rs:5-9);  unacquired rows are exactly 0.
 *
 * Every entry point returns MRS_OK (0) or a negative MRS_E* status and never aborts the process.
 * When `err` is non-NULL the failure message (the reference's panic text where one exists) is
 * written there, truncated to `errlen` bytes.
 */
#ifndef MRSIM_GPU_H
#define MRSIM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRS_OK 0
#define MRS_E_INVALID (-1)   /* bad shape / value: the reference's assert! or validator Err */
#define MRS_E_DEVICE (-2)    /* a HIP call failed */
#define MRS_E_NOMEM (-3)     /* device or host allocation failed */
#define MRS_E_UNSUPPORTED (-4)

#define MRS_MAX_COMP 4       /* compartments per unit the kernels are instantiated for */
#define MRS_MAX_COILS 64

/* Complex f64, the reference's `C { re, im }` (src/kspace.rs:29-33). */
typedef struct { double re, im; } mrs_c64;
/* Complex f32: the device output format (numpy/torch complex64 layout). */
typedef struct { float re, im; } mrs_c32;

/* PartialFourierMode (src/kspace.rs:84-106). */
enum { MRS_PF_FIBERFOX = 0, MRS_PF_CONTIGUOUS = 1, MRS_PF_SCANNER = 2 };
/* EchoFormation (src/kspace.rs:205-209). */
enum { MRS_ECHO_SPIN = 0, MRS_ECHO_GRADIENT = 1 };
/* KspaceWindow (src/kspace.rs:126-131): Tukey uses window_a = alpha; Fermi window_a = radius,
 * window_b = width. */
enum { MRS_WINDOW_NONE = 0, MRS_WINDOW_TUKEY = 1, MRS_WINDOW_HANN = 2, MRS_WINDOW_FERMI = 3 };

/* Acquisition (src/kspace.rs:166-197), every field in the reference's units (ms, Hz, rad).
 * mrs_acquisition_default() gives `Acquisition::default()` (src/kspace.rs:239-267). */
typedef struct {
    double t_line;          /* ms per PE line */
    double t_echo;          /* ms */
    double t_inhom;         /* ms, T2' used when a slice gives no t_inhom */
    double signal_scale;
    int32_t reverse_phase;
    int32_t do_distortions;
    int32_t do_relaxation;
    int32_t pf_mode;        /* MRS_PF_* */
    double noise_variance;
    double partial_fourier;
    double ghost_offset;
    double eddy_strength;
    double eddy_quad;
    double eddy_phase;
    double eddy_tau;        /* ms */
    double spike_amplitude;
    uint64_t n_spikes;
    int32_t window;         /* MRS_WINDOW_* */
    int32_t echo;           /* MRS_ECHO_* */
    double window_a, window_b;
    uint64_t n_coils;
    uint64_t accel;
    uint64_t acs_lines;
    uint64_t seed;
} mrs_acquisition;

/* T2Slice (src/kspace.rs:380-384): `map` NULL means Uniform(value), else Map(map[snx*sny]). */
typedef struct {
    const float* map;
    float value;
} mrs_t2_slice;

/* SliceInput (src/kspace.rs:458-489).  Host pointers.  `t_inhom` and `phase0` may be NULL
 * (the reference's None).  `compartments[q]` and the maps are snx*sny, layout x + snx*y. */
typedef struct {
    const float* const* compartments;
    size_t n_compartments;
    const mrs_t2_slice* t2;        /* n_compartments entries */
    const mrs_t2_slice* t_inhom;   /* n_compartments entries, or NULL */
    const float* fmap;             /* Hz */
    const double* phase0;          /* rad, or NULL */
    size_t sim[2];                 /* [snx, sny] */
    size_t acq_matrix[2];          /* [nx, ny] */
    size_t z, nz;
    int32_t has_eddy_drive;        /* eddy_drive: Option<[f64;3]> */
    double eddy_drive[3];
    int32_t has_prep_drive;        /* informational at slice level, as in the reference */
    double prep_magnitude;
    double prep_direction[3];
    uint64_t slice_seed;
} mrs_slice_input;

/* LineTiming (src/kspace.rs:291-296): ny entries each. */
typedef struct {
    const double* t_ms;
    const double* trf_ms;
    const double* tread_ms;
    const int8_t* polarity;
    size_t ny;
} mrs_line_timing;

/* ---------------------------------------------------------------- host-side protocol tables */

/* Acquisition::default() (src/kspace.rs:239-267). */
int mrs_acquisition_default(mrs_acquisition* acq);

/* LineTiming::for_acquisition (src/kspace.rs:310-326) built from SingleShotEpi
 * (src/readout.rs:45-69) and line_times (src/kspace.rs:271-284).  Each output holds ny values. */
int mrs_line_timing_for_acquisition(const mrs_acquisition* acq, size_t nx, size_t ny, double* t_ms,
                                    double* trf_ms, double* tread_ms, int8_t* polarity);

/* sampling_mask (src/kspace.rs:546-588): nx*ny flags (1 = acquired), layout kx + nx*ky. */
int mrs_sampling_mask(size_t nx, size_t ny, const mrs_acquisition* acq, uint8_t* mask);

/* coil_sensitivity (src/kspace.rs:226-237) evaluated on the simulation grid with the forward
 * model's half-cell registration (src/kspace.rs:701-702), times `scale`, for coils
 * [coil0, coil0 + n) of `ncoils`: out[c*snx*sny + x + snx*y], imaginary parts 0. */
int mrs_ring_coil_maps(size_t coil0, size_t n, size_t ncoils, size_t snx, size_t sny, size_t nx, size_t ny,
                       double scale, mrs_c32* out);

/* ---------------------------------------------------------------- drop-in per-slice forward */

/* build_coil_kspace_timed (src/kspace.rs:607-1042): one coil's acquired k-space, nx*ny complex
 * f64 in `out` (layout kx + nx*ky), normalised by 1/(snx*sny), with spikes and thermal noise on
 * acquired samples.  `timing` NULL uses LineTiming::for_acquisition; `log_amp` may be NULL.
 * Host buffers; the sums run on the current HIP device in FP32 (split-BF16 MFMA) with FP64
 * range-reduced phases. */
int mrs_build_coil_kspace(const mrs_slice_input* inp, const mrs_acquisition* acq, size_t coil, size_t ncoils,
                          const mrs_line_timing* timing, const double* log_amp, mrs_c64* out, char* err,
                          size_t errlen);

/* build_coil_kspaces (src/kspace.rs:1248-1251): all acq->n_coils coils of one slice in one launch,
 * out[coil*nx*ny + kx + nx*ky].  Same semantics as calling mrs_build_coil_kspace per coil. */
int mrs_build_coil_kspaces(const mrs_slice_input* inp, const mrs_acquisition* acq, const mrs_line_timing* timing,
                           const double* log_amp, mrs_c64* out, char* err, size_t errlen);

/* ---------------------------------------------------------------- batched plan (device buffers)
 * A plan is used by one host thread at a time (its calls read and update the plan's host state);
 * distinct plans may be used from different threads.  Work is enqueued on the caller's stream. */

/* Arithmetic of the contractions.  Phases are formed and range-reduced in FP64 in both. */
enum {
    MRS_PREC_SPLIT_BF16 = 0, /* three-term split-BF16 MFMA products, FP32 accumulation */
    MRS_PREC_FP32 = 1        /* plain FP32 VALU contractions, one launch per unit ("plain") */
};

/* Geometry and protocol shared by a group of units. */
typedef struct {
    size_t snx, sny, nx, ny;
    size_t n_coils;                  /* coils per unit (rows of the output) */
    size_t n_compartments;           /* 1..MRS_MAX_COMP */
    int32_t complex_images;          /* 0: real compartments + phase0; 1: complex64 images */
    mrs_acquisition acq;
    const mrs_line_timing* timing;   /* host table, NULL = mrs_line_timing_for_acquisition */
} mrs_group_desc;

/* Per-slice static inputs (device pointers), shared by every unit that names the slice. */
typedef struct {
    int32_t group;
    const float* fmap;               /* P = snx*sny, Hz */
    const double* phase0;            /* P rad, or NULL */
    const float* t2;                 /* Q*P maps (ms) or NULL: t2_uniform[q] */
    const float* t_inhom;            /* Q*P maps (ms) or NULL: t_inhom_uniform[q] if has_t_inhom,
                                        else acq.t_inhom */
    const mrs_c32* coil_maps;        /* n_coils*P complex sensitivities (signal_scale applied by
                                        the plan) */
    const double* log_amp;           /* P, or NULL */
    float t2_uniform[MRS_MAX_COMP];
    float t_inhom_uniform[MRS_MAX_COMP];
    int32_t has_t_inhom;
    size_t z, nz;
    int32_t has_eddy_drive;
    double eddy_drive[3];
} mrs_slice_desc;

/* One unit = one (slice, volume, echo) forward: n_coils x ny x nx complex64 out (unacquired rows
 * zero).  `images`: Q planes of P float (real) or P complex64 (complex_images).  Device pointers. */
typedef struct {
    int32_t slice;
    const void* images;
    mrs_c32* out;
    uint64_t slice_seed;             /* spikes and noise streams (src/kspace.rs:1013, 1031) */
    uint64_t coil0;                  /* global index of the unit's first coil, for noise seeds */
} mrs_unit_desc;

typedef struct mrs_plan mrs_plan;

/* Validate, precompute tables (line tables, twiddles) and allocate device scratch.  Pointers in
 * `slices` and `units` are read on every execute; update them with mrs_plan_set_pointers. */
int mrs_plan_create(const mrs_group_desc* groups, size_t ngroups, const mrs_slice_desc* slices, size_t nslices,
                    const mrs_unit_desc* units, size_t nunits, int32_t precision, mrs_plan** plan, char* err,
                    size_t errlen);
int mrs_plan_set_pointers(mrs_plan* plan, const mrs_slice_desc* slices, const mrs_unit_desc* units, char* err,
                          size_t errlen);
/* Forward on `stream` (a hipStream_t; NULL = the null stream).  Asynchronous. */
int mrs_plan_forward(mrs_plan* plan, void* stream, char* err, size_t errlen);
/* Record the forward into a HIP graph once and replay it on later calls (same pointers). */
int mrs_plan_forward_graph(mrs_plan* plan, void* stream, char* err, size_t errlen);

/* Adjoint (vector-Jacobian product) of the noiseless forward.  Per unit u:
 *   grad_out[u]     n_coils*ny*nx complex64 cotangent (PyTorch's conjugate-Wirtinger gradient)
 *   grad_images[u]  Q planes of P complex64 (complex_images) or P float (real images), or NULL
 *   grad_fmap[u]    P float (Hz), or NULL
 *   grad_t_inhom[u] Q*P float (ms), or NULL
 * Gradients are written (not accumulated). */
int mrs_plan_vjp(mrs_plan* plan, const mrs_c32* const* grad_out, void* const* grad_images, float* const* grad_fmap,
                 float* const* grad_t_inhom, void* stream, char* err, size_t errlen);
/* mrs_plan_vjp for slice inputs unchanged (same pointers, same contents) since this plan's last
 * mrs_plan_forward / mrs_plan_forward_graph on the same stream: the adjoint reads the slice planes
 * (phase, decay rates, coil planes) that forward built instead of rebuilding them, as a backward pass
 * reads saved activations.  Rebuilds them when no forward has run since a slice pointer changed.
 * The caller vouches for the contents. */
int mrs_plan_vjp_after_forward(mrs_plan* plan, const mrs_c32* const* grad_out, void* const* grad_images,
                               float* const* grad_fmap, float* const* grad_t_inhom, void* stream, char* err,
                               size_t errlen);
/* Average duration (ms) of each forward stage over `iters` eager launches, timed with HIP events
 * on `stream`: ms[0] slice prep, ms[1] y-stage, ms[2] x-stage, ms[3] zero rows + spikes/noise,
 * ms[4] the attached partition transform (mrs_plan_set_zstage). */
int mrs_plan_stage_times(mrs_plan* plan, void* stream, int32_t iters, double* ms, char* err, size_t errlen);
/* Terms of the literal sum per batch (sum of n_coils*nx*Ly*snx*sny) and acquired lines. */
int mrs_plan_stats(const mrs_plan* plan, double* terms, double* lines, double* units);
/* Human-readable layout of a plan (groups, lines, chain passes, coil-plane sharing, tiles).  No
 * counterpart in the reference: an inspection aid for callers tuning batches. */
int mrs_plan_describe(const mrs_plan* plan, char* out, size_t outlen);
int mrs_plan_destroy(mrs_plan* plan);

/* Kernel tile parameters chosen by the autotuner: every candidate's graph replay timed on `stream`,
 * the six fastest again round-robin, the lowest median kept (the plan's graph is rebuilt per
 * candidate).  `report` lists the times. */
int mrs_plan_tune(mrs_plan* plan, void* stream, int32_t trials, char* report, size_t reportlen);

/* Set / read the y-stage tiles directly: x columns per block (1..64), 16-line tiles per pass
 * (1, 2 or 4) and units per work item (1, 2 or 4: units of one slice evaluated together, sharing
 * the phase and decay; above 1 only with 4 line tiles, and used by classes of up to 8 coils
 * without the eddy term). */
int mrs_plan_set_tiles(mrs_plan* plan, int32_t x_tile, int32_t line_tiles, int32_t units_per_item, char* err,
                       size_t errlen);
int mrs_plan_get_tiles(const mrs_plan* plan, int32_t* x_tile, int32_t* line_tiles, int32_t* units_per_item);

/* Slice-basis y-stage on (1, the default) or off (0).  With it on, the units of a slice that has
 * four or more of them (its volumes; up to 8 coils, no eddy polynomial, at most 64 acquired lines
 * and 256 padded sim rows) run in one kernel that generates the slice's per-line relaxation and
 * phase factors once for all of them and puts every unit's coil-weighted images on the MFMA (the
 * per-volume loop of acquire_volumes_complex::per_vol, src/kspace.rs:2012-2030, around
 * build_coil_kspace_timed, :607-1042).  Other units keep the per-unit y-stage.  get_basis reports
 * the setting and the number of eligible slices. */
int mrs_plan_set_basis(mrs_plan* plan, int32_t on, char* err, size_t errlen);
int mrs_plan_get_basis(const mrs_plan* plan, int32_t* on, int32_t* eligible_slices);

/* Slice-basis y-adjoint of mrs_plan_vjp on (1, the default) or off (0).  With it on, the units of a
 * slice that qualifies for the slice-basis y-stage (same conditions) take their y-adjoint in one
 * kernel per (set of up to 20 units, column) that generates the slice's per-line phase and decay
 * factors once for all of them and contracts every unit's x-adjoint result against them over the
 * lines on MFMA.  Split-BF16 plans only; other units keep the per-unit y-adjoint.  No counterpart in
 * the reference (which has no adjoint): a planner control of this build's vector-Jacobian product. */
int mrs_plan_set_vjp_basis(mrs_plan* plan, int32_t on, char* err, size_t errlen);
int mrs_plan_get_vjp_basis(const mrs_plan* plan, int32_t* on);

/* ---- 3D gradient-echo train (stack of EPI blocks over nz partitions) ----------------------- */

/* Within-echo timing of a GRASE/3D-GE block on ny lines: t_ms[ky] from the echo centre and the
 * readout polarity (+1/-1, alternating with the acquisition index within the block).
 * Replaces readout::grase_block (src/readout.rs:206-221); refuses ny not divisible by
 * ky_segments with the reference's text. */
int mrs_grase_block(size_t ny, size_t ky_segments, double t_line_ms, int32_t reverse_phase, double* t_ms,
                    int8_t* polarity, char* err, size_t errlen);

/* Line table of a 3D gradient-echo train on ny x nz (index p*ny + ky): the shot (sy*kz_segments + sz)
 * and the excitation within it that reads each line.  kz_order 0 = linear, 1 = centric.  Any output
 * may be NULL.  Replaces readout::ge3d_lines (src/readout.rs:428-462) with readout::kz_order
 * (src/readout.rs:170-189). */
int mrs_ge3d_lines(size_t ny, size_t nz, size_t ky_segments, double t_line_ms, int32_t reverse_phase,
                   size_t kz_segments, int32_t kz_order, int32_t* shot, int32_t* excitation, size_t* n_shots,
                   size_t* n_exc, size_t* e_c, char* err, size_t errlen);

/* Timing checks of a 3D gradient-echo train (echo blocks clear of the excitation pulse, of each other,
 * of the next excitation and of the repetition).  Replaces readout::check_ge3d_timing
 * (src/readout.rs:465-501); the refusal texts are the reference's. */
int mrs_check_ge3d_timing(const double* echo_times_ms, size_t n_echoes, double exc_spacing_ms,
                          double excitation_time_ms, size_t ny, size_t ky_segments, double t_line_ms,
                          int32_t reverse_phase, size_t n_exc, double t_exc_ms, double tr_ms, char* err,
                          size_t errlen);

/* One set of the batched partition transform: nz partitions of n_coils*ny*nx complex64 samples,
 * layout [p][coil][ky][kx] (device pointers).
 *   forward: out[p] (+)= sum_c w(p, ky % ky_segments, c) sum_z zk[p][z] in_c[z]
 *   adjoint: out_c[z] (+)= sum_p conj(zk[p][z]) w(p, ky % ky_segments, c) in[p]
 * zk[p][z] = exp(+i 2 pi (p - nz/2)(z - nz/2)/nz); in_c (forward) or out_c (adjoint) is compartment c's
 * block, comp_stride complex elements apart; weights[(p*ky_segments + sy)*n_comp + c] (NULL = 1).
 * Replaces the z-DFT and weights of kspace3d::GePlan::{compartments,kspace}
 * (src/kspace3d.rs:677-692, 716-722) with zkernel (src/kspace3d.rs:90-100).  nz <= 64. */
typedef struct mrs_zset {
    const mrs_c32* in;
    mrs_c32* out;
    const float* weights;
    size_t comp_stride;
    uint32_t nz, n_comp, n_coils, ny, nx, ky_segments;
    int32_t accumulate;
    int32_t reserved;
} mrs_zset;

/* Run sets on `stream` and wait for them (the call allocates and frees its descriptor copy). */
int mrs_zstage(const mrs_zset* sets, size_t nsets, int32_t adjoint, void* stream, char* err, size_t errlen);

/* The same partition transform with its descriptors kept on the device, for a caller that runs one
 * set list many times (the adjoint of a 3D train, src/kspace3d.rs:90-100, 716-722): create uploads
 * them once (blocking), run enqueues the sets on `stream` and returns at once, destroy frees them.
 * The sets' buffers must stay allocated while the z-plan runs. */
typedef struct mrs_zplan mrs_zplan;
int mrs_zplan_create(const mrs_zset* sets, size_t nsets, int32_t adjoint, mrs_zplan** zplan, char* err, size_t errlen);
int mrs_zplan_run(mrs_zplan* zplan, void* stream, char* err, size_t errlen);
void mrs_zplan_destroy(mrs_zplan* zplan);

/* Spikes and receiver noise of one volume's acquired 3D k-space (device, nz x n_coils x ny x nx complex64,
 * in place): partitions_to_slices (src/kspace3d.rs:235-264) before its inverse z-DFT.  `volume` is the
 * volume index g, `seed` the echo's seed (series seed ^ echo_salt(e)); acq gives n_spikes,
 * spike_amplitude, noise_variance and the sampling mask.  The inverse z-DFT is mrs_zstage's adjoint
 * with weights 1/nz.  Waits for `stream`. */
int mrs_partition_noise(mrs_c32* kspace, size_t nz, size_t n_coils, size_t nx, size_t ny, const mrs_acquisition* acq,
                        uint64_t volume, uint64_t seed, void* stream, char* err, size_t errlen);

/* Attach forward sets to a plan: they run after the 2D forward inside mrs_plan_forward and its graph.
 * nsets = 0 detaches.  Sets that accumulate into the same output must not be attached together. */
int mrs_plan_set_zstage(mrs_plan* plan, const mrs_zset* sets, size_t nsets, char* err, size_t errlen);

/* ---- reconstruction (the observation point of the reference's golden fixtures) -------------- */

/* Batched reconstruction of one geometry: per unit, reconstruct_coils_full (src/kspace.rs:1301-1378)
 * without GRAPPA -- the window (KspaceWindow::at, :140-161) on every sample, the centred inverse 2D
 * DFT of each coil (inverse_2d, :1428-1453), the phase-preserving Roemer combine
 * sum_c s_c img_c / max(sum_c s_c^2, 1e-12) with the real sensitivities `sensitivities`
 * (n_coils x ny x nx, x + nx*y; NULL = coil_sensitivity on the acquired grid, :1353-1355).
 * n_coils must be max(acq->n_coils, 1) ("one k-space per coil").  Under acceleration the caller applies
 * GRAPPA to the k-spaces first (mrs_grappa_apply).  nx, ny <= 128.  The contractions run as
 * split-BF16 MFMA with FP32 accumulation. */
typedef struct mrs_recon mrs_recon;
int mrs_recon_create(size_t nx, size_t ny, size_t n_coils, const mrs_acquisition* acq, const float* sensitivities,
                     mrs_recon** recon, char* err, size_t errlen);
/* kspace[u]: n_coils x ny x nx complex64 (kx + nx*ky per coil); image[u]: ny x nx complex64.
 * Device pointers; runs on `stream` and waits for it. */
int mrs_recon_run(mrs_recon* recon, const mrs_c32* const* kspace, mrs_c32* const* image, size_t n, void* stream,
                  char* err, size_t errlen);
/* The same in two steps: upload the unit list once, then launch it asynchronously on `stream`
 * (no allocation or copy: a HIP graph can record the launch). */
int mrs_recon_set_units(mrs_recon* recon, const mrs_c32* const* kspace, mrs_c32* const* image, size_t n, char* err,
                        size_t errlen);
int mrs_recon_launch(mrs_recon* recon, void* stream, char* err, size_t errlen);
int mrs_recon_destroy(mrs_recon* recon);

/* GRAPPA (src/kspace.rs:1498-1633), in FP64 on the GPU.  Weights: (accel-1) x n_coils x 6*n_coils complex
 * f64, weights[d-1][output coil][feature], feature (line, point, coil) = 2 source lines x readout
 * offsets {-1, 0, +1} (wrapped) x coils.
 * mrs_grappa_fit: GrappaKernel::fit (:1540-1598) on a fully sampled reference scan `ref_kspace`
 * (device, n_coils x ny x nx complex64) over the acs_lines about ny/2; `lines` (host, ny flags) are
 * the PE lines the reference scan read; `weights` is device memory.  Tikhonov 1e-3 of the mean
 * diagonal; solve_hermitian (:1457-1491).  Waits for `stream`. */
int mrs_grappa_fit(const mrs_c32* ref_kspace, size_t n_coils, size_t nx, size_t ny, size_t accel, size_t acs_lines,
                   const uint8_t* lines, mrs_c64* weights, void* stream, char* err, size_t errlen);
/* mrs_grappa_fit without allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results as mrs_grappa_fit, but the fit's scratch (A^H A and
 * A^H b of every offset, 16 * (accel-1) * (36 n_coils^2 + 6 n_coils^2) bytes, and the line flags) lives
 * in one library-owned device array of 8 MiB per GPU, kept until the process ends; at most 1024 PE
 * lines.  Larger problems are refused with MRS_E_UNSUPPORTED and a message.  Calls that overlap on one
 * GPU share that array and race: serialise them (one stream, or one caller at a time).  Waits for
 * `stream`. */
int mrs_grappa_fit_noalloc(const mrs_c32* ref_kspace, size_t n_coils, size_t nx, size_t ny, size_t accel,
                           size_t acs_lines, const uint8_t* lines, mrs_c64* weights, void* stream, char* err,
                           size_t errlen);
/* GrappaKernel::apply (:1605-1633): synthesize every missing PE line of n units' undersampled coil
 * k-spaces (device pointers, n_coils x ny x nx complex64) in place; `acquired` (host, ny flags) are
 * the lines the units read.  Waits for `stream`. */
int mrs_grappa_apply(mrs_c32* const* kspace, size_t n, size_t n_coils, size_t nx, size_t ny, size_t accel,
                     const uint8_t* acquired, const mrs_c64* weights, void* stream, char* err, size_t errlen);

/* reconstruct_coils (src/kspace.rs:1290-1298) for host buffers: coil_kspace n_coils x (kx + nx*ky)
 * complex f64 -> out (x + nx*y) complex f64, on the GPU.  Under acceleration `grappa_weights`
 * (host, as mrs_grappa_fit writes them) is required and applied first. */
int mrs_reconstruct_coils(const mrs_c64* coil_kspace, size_t n_coils, size_t nx, size_t ny,
                          const mrs_acquisition* acq, const mrs_c64* grappa_weights, mrs_c64* out, char* err,
                          size_t errlen);

/* simulate_slice (src/kspace.rs:1229-1235, 1273-1284) entirely on the GPU: every coil's k-space,
 * under acceleration the GRAPPA kernel fitted on the slice's own reference scan (slice seed salted by
 * 0x4752_5043_414C, no eddy or preparation drive, no acceleration, no spikes) and applied, then the
 * reconstruction.  out: nx*ny interleaved (re, im) f32, the reference's Vec<(f32, f32)>. */
int mrs_simulate_slice(const mrs_slice_input* inp, const mrs_acquisition* acq, float* out, char* err, size_t errlen);

/* The tail of acquire_volumes (src/kspace.rs:2036-2060, 1840-1846) on a series' reconstructed images:
 * images (device) [g][vox] complex64, vox = x + nx*(y + ny*z); noise_sigma (device, nvox, or NULL):
 * image-space complex Gaussian noise of that SD per component from
 * Rng((receiver_seed ^ g*0x9E37_79B9_7F4A_7C15 ^ vox*0xBF58_476D_1CE4_E5B9) | 1); out0/out1 (device,
 * nvox*n_volumes f32) in the entry point's layout vox*n_volumes + g: mode 0 the real and imaginary
 * parts (simulate_acquisition_complex), mode 1 magnitude and phase (simulate_acquisition_oversampled).
 * Waits for `stream`. */
int mrs_series_epilogue(const mrs_c32* images, size_t n_volumes, size_t nvox, const float* noise_sigma,
                        uint64_t receiver_seed, int32_t mode, float* out0, float* out1, void* stream, char* err,
                        size_t errlen);

/* Nufft1 (src/nufft.rs:25-113) on the GPU in FP64: the type-1 NUFFT
 * F[k] = sum_j w_j exp(+i 2 pi k v_j / n), k = k_lo .. k_lo + n_out - 1, by gridding with the
 * exponential-of-semicircle kernel at oversampling 2 (width from `tol` as the reference chooses it).
 * Device buffers: pos [pos_cols][nsrc] cells (column c uses pos[c % pos_cols]); weights
 * [ncols][nw][nsrc]; out [ncols][nw][n_out].  The output band must sit inside [-n/2, n/2]; the fine
 * grids of a column (nw * 2n complex f64) must fit in 160 KiB of LDS.  Waits for `stream`. */
int mrs_nufft1(const double* pos, size_t pos_cols, const mrs_c64* weights, size_t nsrc, size_t nw, size_t ncols,
               size_t n, long k_lo, size_t n_out, double tol, mrs_c64* out, void* stream, char* err, size_t errlen);

/* build_coil_kspaces (src/kspace.rs:1248-1251) as the reference's `kspace` feature build computes it
 * where its NUFFT y-stage applies (src/kspace.rs:778-831): uniform relaxation, no eddy polynomial,
 * ny >= 3, line times affine per parity with the parity's readout polarity; other inputs are refused
 * with the reason.  FP64 on the GPU: the NUFFT y-stage (tolerance 1e-13), the per-line relaxation
 * combine, the twiddle x-stage, spikes and noise.  Host buffers as mrs_build_coil_kspaces; the
 * comparator of the exact forward. */
int mrs_build_coil_kspaces_nufft(const mrs_slice_input* inp, const mrs_acquisition* acq, mrs_c64* out, char* err,
                                 size_t errlen);

/* Size of a public struct as compiled into the library (ABI check for bindings). */
size_t mrs_sizeof(const char* name);

/* Device evaluation of sin/cos(2 pi cyc[i]) (host buffers), to measure the phase arithmetic's
 * error: mode 0 = as the kernels evaluate it, 1 = FP64 quarter-turn reduction + FP32 polynomials,
 * 2 = FP64 reduction + hardware v_sin_f32/v_cos_f32. */
int mrs_probe_sincos(const double* cyc, float* s, float* c, size_t n, int32_t mode, char* err, size_t errlen);

/* Library version string. */
const char* mrs_version(void);

#ifdef __cplusplus
}
#endif
#endif
