// internal.h -- device-side records shared by the HIP kernels (kernels.hip) and the host plan
// (plan.cpp).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/mrsim_gpu.h"

namespace mrs {

// One acquired phase-encode line of a timing group, in the order the kernels walk them: sorted by
// x-stage table (the Nyquist-ghost polarity), then ky; each table segment padded to a multiple of
// 16 lines with ky = -1 records.
struct LineRec {
    double t_s;    // fieldmap phase time t_ms/1000 (s), src/kspace.rs:844, plus TE/1000 for a gradient
                   // echo with distortions (its f*TE phase, :717-721)
    double kyn;    // (ky - ny/2) / sny, src/kspace.rs:729
    double d;      // eddy clock exp(-tread/tau_eddy) * t (s), src/kspace.rs:910; 0 without eddy
    float trf;     // ms from the RF (T2 decay), src/kspace.rs:845
    float tau;     // ms of T2' decay: |t|*1000 (spin echo) or trf (gradient echo), :896-897
    float tb0;     // s of fieldmap phase accrued (t, or t + TE/1000 for gradient echo): d phase/d fmap
    int32_t ky;    // output row, -1 for padding
    int32_t table; // x-stage table 0 (+ghost) or 1 (-ghost), src/kspace.rs:987, 1109
    int32_t cls;   // side of the echo: 0 for t >= 0, 1 for t < 0 (selects the shared decay ratios)
};
static_assert(sizeof(LineRec) == 48, "LineRec layout");

// Static per-slice planes, written by the prep kernel, all in [x][y] (y fastest) order so a lane's
// four consecutive y values are one 16-byte load.
struct DevSlice {
    double* base;      // P: object phase in cycles (phase0 + eddy ramp)/2pi
    float* rate;       // P: fieldmap (Hz) if do_distortions else 0
    double* poly;      // P: eddy polynomial (cycles per second of eddy clock), 0 without eddy
    float* nr2;        // Q*P: -log2(e)/T2 (1/ms), 0 without relaxation
    float* nri;        // Q*P: -log2(e)/T2' (1/ms)
    uint32_t* s_hi;    // [x][c][y] packed bf16 (re | im<<16) of signal_scale*exp(log_amp)*S_c, c padded
                       // with zero coils to RB*8
    uint32_t* s_lo;    // residual of the same after s_hi, bf16
    float2* s32;       // [c][x][y] the same in FP32 (plain path)
    float2* rot;       // [x][y] chain phase rotor cis(2 pi (rate dt + dkyn yy)) when the group shares one
    float* rq;         // [cls][q][x][y] chain decay ratios exp2(dtrf nr2 + dtau_cls nri)
    // user inputs (device)
    const float* fmap;
    const double* phase0;
    const float* t2;
    const float* t_inhom;
    const mrs_c32* coil_maps;
    const double* log_amp;
    float t2_uniform[MRS_MAX_COMP];
    float ti_uniform[MRS_MAX_COMP];
    int32_t has_t_inhom;
    int32_t group;
    int32_t z, nz;
    int32_t has_eddy;
    double eddy[3];
    int32_t s_owner;   // this slice writes the coil planes; otherwise they alias an earlier slice's
    int32_t p_owner;   // this slice writes the static planes (base .. rq); otherwise the same
    int32_t prep_extra;   // plane forms written in addition to the plan's (PREP_S32 for the coil-plane
                          // owner of slices whose volumes run in the slice-basis y-stage)
};

// Slice prep outputs: the split coil planes (MFMA forward) and/or the FP32 ones (plain, adjoint).
enum { PREP_SPLIT = 1, PREP_S32 = 2 };

struct DevGroup {
    int32_t snx, sny, nx, ny;
    int32_t C, Q, RB;          // coils, compartments, 16-row blocks of 2C
    int32_t Ly;                // acquired lines
    int32_t rotor_passes;      // leading 64-line passes whose lanes' 4 lines are arithmetic progressions
    int32_t rot_shared;        // every chain steps by the same (dt, dkyn, dtrf) and, per echo side, dtau
    double rot_dt, rot_dk;     // the shared steps (s, cycles per voxel of yy)
    float rot_dtrf, rot_dtau[2];
    int32_t LyP;               // padded lines (multiple of 16)
    int32_t cplx;              // complex images
    int32_t echo, do_relax, do_dist;
    double sys_yoff;           // yy = y - (sys + yoff)
    double sxs_xoff;
    int32_t ox, oy;
    double t_echo, t_inhom, signal_scale;
    double eddy_strength, eddy_quad, eddy_phase;
    const LineRec* lines;      // LyP records
    const uint32_t* t_hi[2];   // [kx][x] packed bf16 twiddles / (snx*sny)
    const uint32_t* t_lo[2];
    const float2* t32[2];      // [kx][x] FP32 twiddles / (snx*sny)
    const int32_t* zero_rows;  // unacquired ky rows
    int32_t n_zero_rows;
    int32_t noise;             // noise_variance > 0
    int32_t n_spikes;
    double noise_sigma;
    double spike_amplitude;
    // How the slice-basis y-adjoint forms its T2' and fieldmap sums from the line weights tau_l and
    // tb0_l: 0, one product each; 1, tau_l = adj_alpha * tb0_l on every acquired line, so the T2' sum
    // is adj_alpha times the fieldmap sum; 2, tau_l = adj_alpha * |tb0_l|, with the lines in the
    // adjoint's K order (adj_perm) by the sign of tb0 (tb0 >= 0 first, each side padded to 16 lines),
    // so one tb0-weighted product split at K step adj_ksplit gives both sums.
    int32_t adj_mode, adj_nks, adj_ksplit;
    float adj_alpha;
    int8_t adj_perm[64];       // K position -> line (-1: none); modes 0 and 1: the identity up to LyP
};

struct DevUnit {
    const void* images;   // Q planes of P float or complex64 (user layout x + snx*y)
    float2* out;          // C x ny x nx
    float2* gbuf;         // y-stage result [x][n], column n = line*C + coil (plan scratch):
                          // complex f32 in the FP32 path; in the split path each complex is {hi, lo}
                          // of the split of (re, im) (packed re | im<<16 BF16 pairs), the x-stage's
                          // B operand as is
    int32_t slice;
    int32_t group;
    uint64_t seed;
    uint64_t coil0;
};

// A slice-basis set: up to kBasisMax units (volumes) of one slice whose y-stage shares the slice's
// per-line basis; stored as [count, unit 0, ..., unit kBasisMax-1].
static constexpr int kBasisMax = 20;
static constexpr int kBasisStride = kBasisMax + 1;
static constexpr int kBasisMin = 4;      // fewer units of a slice run in the per-unit y-stage
static constexpr int kBasisMaxY = 256;   // largest padded sim-grid height (snyP) of the basis y-stage

// Work item of the batched kernels.
struct WorkItem {
    int32_t unit;
    int32_t tile;   // x tile (y-stage) or column tile (x-stage)
};

// One z-stage set (the batched 3D partition transform): nz partitions of ncols = coils * ny * nx
// complex samples, layout [partition][coil][ky][kx]; n_comp compartment blocks cstride apart on the
// side that carries compartments (the input of the forward, the output of the adjoint).
struct ZSet {
    const float2* in;
    float2* out;
    const float* w;       // [nz][kseg][nq] per-excitation weights, or null (all 1)
    int64_t cstride;      // complex elements between compartment blocks
    int32_t nz, ncols, ny, nx, kseg, nq, accumulate, mt;
};

struct ZBlock {
    int32_t set;
    int32_t col0;
};

// Reconstruction of one geometry (reconstruct_coils_full, src/kspace.rs:1301-1378): tables on the
// device, matrices zero-padded to multiples of 16.
struct ReconGeom {
    int32_t nx, ny, nxP, nyP, C;
    const uint32_t* fy_hi;   // [nyP][nyP] packed bf16 of Fy[y][ky] = cis(-2 pi (ky-ys)(y-ys)/ny)
    const uint32_t* fy_lo;
    const uint32_t* fx_hi;   // [nxP][nxP] packed bf16 of Fx[x][kx] = cis(-2 pi (kx-xs)(x-xs)/nx)
    const uint32_t* fx_lo;
    const float* window;     // [ny][nx] reconstruction window, or null
    const float* sens;       // [C][nyP][nxP] Roemer sensitivities (real)
    const float* rssum;      // [nyP][nxP] 1 / max(sum_c sens^2, 1e-12)
};

struct ReconItem {
    const float2* kspace;    // C x ny x nx complex64
    float2* image;           // ny x nx complex64 (x + nx*y)
};

double ring_sensitivity(size_t coil, size_t n_coils, double x, double y, size_t nx, size_t ny);

int build_zsets(const mrs_zset* in, size_t n, bool adjoint, std::vector<ZSet>& sets, std::vector<ZBlock>& blocks,
                int off[4], char* err, size_t errlen);
hipError_t enqueue_zsets(const ZSet* d_sets, const ZBlock* d_blocks, const int off[4], bool adjoint, hipStream_t st);

struct VjpUnit {
    const float2* gout;   // cotangent C x ny x nx
    float2* gimg;         // Q planes of P (complex) or P float (real: Q planes float)
    float* gfmap;         // P
    float* gti;           // Q*P
    float2* gbar;         // x-adjoint result [x][LyP][C] (plan scratch)
};

}  // namespace mrs
