// kernels.hip -- CDNA4 (gfx950) kernels of the exact k-space DFT forward model.
//
// Per slice unit, coil c, acquired line l (ky), readout sample kx (src/kspace.rs:607-1042):
//   K_c[kx,ky] = sum_x T_tab(l)[kx,x] * sum_y S_c(x,y) * E(x,y,l)
//   E(x,y,l)   = W(x,y,l) * cis(2 pi cyc(x,y,l))
//   W          = sum_q comp_q(x,y) * exp2(trf_l * nr2_q(x,y) + tau_l * nri_q(x,y))     (relaxation)
//   cyc        = base(x,y) + rate(x,y)*t_l + kyn_l*(y - sys - yoff) + poly(x,y)*d_l  (FP64)
// where T_tab is the x-stage twiddle table of the line's ghost polarity divided by snx*sny,
// S_c = signal_scale * exp(log_amp) * sensitivity, and the eddy polynomial's x, y and z factors
// (src/kspace.rs:908-925, 971, 977-984) are one phase term poly*d.  The phase is formed and
// range-reduced in FP64, then evaluated by an FP32 sincos of the reduced argument.
//
// Two arithmetic variants:
//   split-BF16  y- and x-contractions on MFMA (v_mfma_f32_16x16x32_bf16), every operand split
//               into bf16 hi + lo and multiplied as hi*hi + hi*lo + lo*hi with FP32 accumulation;
//               E is generated in registers directly in the MFMA B-operand layout (no LDS copy).
//   plain FP32  one thread per (x, line) summing over y on the VALU, then per (row, kx).
#include <algorithm>

#include "devcommon.h"

namespace mrs {

// ----------------------------------------------------------------------------- slice prep
// Static per-slice planes (src/kspace.rs:678-725, 908-925).  Block = one 32 x 32 voxel tile of one
// slice: user planes (x fastest) are read coalesced along x, staged in LDS, and written in the
// kernels' [x][y] order coalesced along y.  grid = (tiles_x * tiles_y, listed slices): only slices
// that own a plane set are listed (the others read their owner's planes).
// QM: the most compartments of any slice in the launch (sizes the per-compartment tiles, so that the
// LDS of a block stays within a third of a CU's for up to three compartments).  The static tiles
// (base, poly) and the coil tiles of a pair of coils share their LDS: the coil phase starts after the
// static planes are written.
template <int QM>
__global__ __launch_bounds__(256) void prep_slices(DevSlice* __restrict__ slices, const DevGroup* __restrict__ groups,
                                                   const int32_t* __restrict__ list, int nlist, int mode) {
    __shared__ union {
        double sd[2][32][33];
        float2 sc[2][32][33];
    } lds_u;
    __shared__ float sf[1 + 2 * QM][32][33];
    __shared__ float samp[32][33];
    auto& sd = lds_u.sd;
    auto& sc = lds_u.sc;
    if ((int)blockIdx.y >= nlist) return;
    const int s = list[blockIdx.y];
    DevSlice& sl = slices[s];
    const DevGroup& gr = groups[sl.group];
    const int snx = gr.snx, sny = gr.sny;
    const int snxP = (snx + 15) & ~15, snyP = (sny + 15) & ~15;
    const int tiles_x = (snxP + 31) / 32, tiles_y = (snyP + 31) / 32;
    if ((int)blockIdx.x >= tiles_x * tiles_y) return;
    const int x0 = (blockIdx.x % tiles_x) * 32, y0 = (blockIdx.x / tiles_x) * 32;
    const int C = gr.C, Q = min(gr.Q, QM);   // the launch's QM covers every slice's Q
    const bool p_own = sl.p_owner != 0;   // read here, where its latency hides behind the read phase
    const size_t PP = (size_t)snxP * snyP, P = (size_t)snx * sny;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // read phase: x fastest
    const double zc = (double)sl.z - (double)(sl.nz / 2);
    const double zc_c = (double)sl.z - ((double)sl.nz - 1.0) / 2.0;
    const double g0 = sl.eddy[0], g1 = sl.eddy[1], g2 = sl.eddy[2];
    const bool eddy_ph = gr.eddy_phase != 0.0 && sl.has_eddy, eddy = gr.eddy_strength != 0.0 && sl.has_eddy;
    for (int k = 0; k < 4; k++) {
        const int yl = ty + 8 * k, x = x0 + tx, y = y0 + yl;
        const bool valid = x < snx && y < sny;
        const size_t i = (size_t)x + (size_t)snx * y;
        double base = 0.0, poly = 0.0;
        float rate = 0.0f, amp = (float)gr.signal_scale;
        if (valid) {
            const double xc = ((double)x - gr.sxs_xoff) / (double)gr.ox;
            const double yc = ((double)y - gr.sys_yoff) / (double)gr.oy;
            double p0 = sl.phase0 ? gld(sl.phase0)[i] : 0.0;
            if (eddy_ph) p0 += gr.eddy_phase * (g0 * xc + g1 * yc + g2 * zc_c);
            const float f = gld(sl.fmap)[i];
            base = p0 / kTau;   // a gradient echo's f*TE/1000 is on the lines' times
            rate = gr.do_dist ? f : 0.0f;
            if (eddy)
                poly = gr.eddy_strength * (g0 * xc + g1 * yc + g2 * zc) +
                       gr.eddy_quad * (g0 * xc * xc + g1 * yc * yc + g2 * zc * zc);
            if (sl.log_amp) amp = (float)(gr.signal_scale * exp(gld(sl.log_amp)[i]));
        }
        sd[0][tx][yl] = base;
        sd[1][tx][yl] = poly;
        sf[0][tx][yl] = rate;
        samp[tx][yl] = amp;
        for (int q = 0; q < Q; q++) {
            float a = 0.0f, b = 0.0f;
            if (valid && gr.do_relax) {
                const float t2 = sl.t2 ? gld(sl.t2)[q * P + i] : sl.t2_uniform[q];
                const float ti = sl.t_inhom ? gld(sl.t_inhom)[q * P + i]
                                            : (sl.has_t_inhom ? sl.ti_uniform[q] : (float)gr.t_inhom);
                a = -1.4426950408889634f / t2;
                b = -1.4426950408889634f / ti;
            }
            sf[1 + 2 * q][tx][yl] = a;
            sf[2 + 2 * q][tx][yl] = b;
        }
    }
    __syncthreads();
    const int wy = threadIdx.x & 31, wx = threadIdx.x >> 5;   // write phase: y fastest
    // static planes: written once per set of slices that would write identical ones
    if (p_own) {
        for (int k = 0; k < 4; k++) {
            const int xl = wx + 8 * k, x = x0 + xl, y = y0 + wy;
            if (x >= snxP || y >= snyP) continue;
            const size_t o = (size_t)x * snyP + y;
            gst(sl.base)[o] = sd[0][xl][wy];
            gst(sl.poly)[o] = sd[1][xl][wy];
            const float rate = sf[0][xl][wy];
            gst(sl.rate)[o] = rate;
            for (int q = 0; q < Q; q++) {
                gst(sl.nr2)[q * PP + o] = sf[1 + 2 * q][xl][wy];
                gst(sl.nri)[q * PP + o] = sf[2 + 2 * q][xl][wy];
            }
            if (gr.rot_shared) {
                // chain rotors shared by every lane of a chain pass (FP64 angle, FP64 reduction)
                const double yy = (double)y - gr.sys_yoff;
                float rs, rc;
                sincos_turns(__builtin_fma(gr.rot_dk, yy, (double)rate * gr.rot_dt), rs, rc);
                reinterpret_cast<GAS f32x2*>(gst(sl.rot))[o] = f32x2{rc, rs};
                for (int c = 0; c < 2; c++)
                    for (int q = 0; q < Q; q++)
                        gst(sl.rq)[(c * Q + q) * PP + o] = __builtin_amdgcn_exp2f(
                            __builtin_fmaf(gr.rot_dtrf, sf[1 + 2 * q][xl][wy], gr.rot_dtau[c] * sf[2 + 2 * q][xl][wy]));
            }
        }
    }
    // coil planes: written once per distinct (maps, amplitude) set, only in the forms asked for
    if (!sl.s_owner) return;
    const int smode = mode | sl.prep_extra;
    const bool w_split = smode & PREP_SPLIT, w_s32 = smode & PREP_S32;
    const int Cp = w_split ? gr.RB * 8 : (w_s32 ? C : 0);   // split maps padded with zero coils to 8-coil row blocks
    // Coils in groups of CG: every coil map value of the group is loaded before the first is
    // written, so a block waits for one memory latency per group rather than one per coil.
    constexpr int CG = 4;
    for (int c0 = 0; c0 < Cp; c0 += CG) {
        f32x2 m[CG][4];
#pragma unroll
        for (int cc = 0; cc < CG; cc++) {
            const int c = c0 + cc;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int yl = ty + 8 * k, x = x0 + tx, y = y0 + yl;
                m[cc][k] = f32x2{0.f, 0.f};
                if (c < C && x < snx && y < sny)
                    m[cc][k] = *at32<f32x2>(sl.coil_maps, (uint32_t)c * (uint32_t)P + (uint32_t)(x + snx * y));
            }
        }
        // two coils per transpose through LDS (one pair of barriers)
#pragma unroll
        for (int cc = 0; cc < CG; cc += 2) {
            const int c = c0 + cc;
            if (c >= Cp) break;
            __syncthreads();   // the previous pair's (or the static planes') reads of this LDS are done
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int yl = ty + 8 * k;
                    const float a = samp[tx][yl];
                    sc[h][tx][yl] = make_float2(m[cc + h][k].x * a, m[cc + h][k].y * a);
                }
            __syncthreads();
            for (int k = 0; k < 4; k++) {
                const int xl = wx + 8 * k, x = x0 + xl, y = y0 + wy;
                if (x >= snxP || y >= snyP) continue;
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int ch = c + h;
                    if (ch >= Cp) break;
                    const float2 v = sc[h][xl][wy];
                    if (w_split) {
                        uint32_t hi, lo;
                        split2(v.x, v.y, hi, lo);
                        const size_t so = ((size_t)x * Cp + ch) * snyP + y;
                        gst(sl.s_hi)[so] = hi;
                        gst(sl.s_lo)[so] = lo;
                    }
                    if (w_s32 && ch < C)
                        reinterpret_cast<GAS f32x2*>(gst(sl.s32))[((size_t)ch * snxP + x) * snyP + y] = f32x2{v.x, v.y};
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------- E generation
// The four voxels (x, y0..y0+3) a lane evaluates for every line it owns, with the images of the V
// units (volumes of one slice) that share every other per-voxel quantity.
template <int Q, bool CPLX, bool EDDY, int V>
struct Vox4 {
    double base[4];
    double poly[EDDY ? 4 : 1];
    float rate[4];
    float nr2[Q][4], nri[Q][4];
    float cr[V][Q][4];
    float ci[V][CPLX ? Q : 1][4];
};

// Phase (sin, cos of 2 pi cyc) and per-compartment decay of voxel (x, y0+j) at one line:
// yy = y0 + j - (sys + yoff).  Shared by every unit of the slice.
template <int Q, bool CPLX, bool EDDY, int V>
__device__ __forceinline__ void phase_decay(const Vox4<Q, CPLX, EDDY, V>& v, int j, double yy, double t_s, double kyn,
                                            double d, float trf, float tau, float& sn, float& cs, float (&dq)[Q]) {
    double cyc = __builtin_fma(kyn, yy, v.base[j]);
    cyc = __builtin_fma((double)v.rate[j], t_s, cyc);
    if constexpr (EDDY) cyc = __builtin_fma(v.poly[j], d, cyc);
    sincos_turns(cyc, sn, cs);
#pragma unroll
    for (int q = 0; q < Q; q++) dq[q] = __builtin_amdgcn_exp2f(__builtin_fmaf(trf, v.nr2[q][j], tau * v.nri[q][j]));
}

// E of unit vv at voxel j: (sum_q m_q dq_q) (cs + i sn).
template <int Q, bool CPLX, bool EDDY, int V>
__device__ __forceinline__ void unit_e(const Vox4<Q, CPLX, EDDY, V>& v, int vv, int j, float sn, float cs,
                                       const float (&dq)[Q], float& er, float& ei) {
    float wr = v.cr[vv][0][j] * dq[0], wi = CPLX ? v.ci[vv][0][j] * dq[0] : 0.0f;
#pragma unroll
    for (int q = 1; q < Q; q++) {
        wr = __builtin_fmaf(v.cr[vv][q][j], dq[q], wr);
        if constexpr (CPLX) wi = __builtin_fmaf(v.ci[vv][q][j], dq[q], wi);
    }
    if constexpr (CPLX) {
        er = wr * cs - wi * sn;
        ei = wr * sn + wi * cs;
    } else {
        er = wr * cs;
        ei = wr * sn;
    }
}

// ----------------------------------------------------------------------------- y-stage operand staging
// A wave's operands of one 16-row chunk (column x, rows ch..ch+15) are copied from HBM/L2 straight
// into a per-wave LDS buffer by direct-to-LDS loads (global_load_lds), one instruction per array:
// SGPR base + a lane-constant 32-bit offset, lane k writing the k-th 16-byte (or 4-byte) piece.  The
// next chunk's copy is in flight while the current chunk is evaluated from the other buffer, so the
// evaluation reads its operands at LDS latency.  Layout of one buffer, in 16-byte pieces:
//   base[16] f64 | poly[16] f64 (eddy) | rate[16] | nr2[Q][16] | nri[Q][16] | s_hi[8RB][16] |
//   s_lo[8RB][16] | rot[16] f32x2 (chains) | rq[2][Q][16] (chains) | images[V][Q][16] (x2 complex)
// The coil planes are staged up to 32 coils (SA); above that their double buffer would cost more
// LDS than two blocks per CU allow, and they are loaded directly.
template <int Q, int RB, bool CPLX, bool EDDY, int V, bool ROT>
struct YStage {
    static constexpr bool SA = RB <= 4;
    static constexpr int O_BASE = 0;
    static constexpr int O_POLY = 8;
    static constexpr int O_RATE = O_POLY + (EDDY ? 8 : 0);
    static constexpr int O_NR2 = O_RATE + 4;
    static constexpr int O_NRI = O_NR2 + 4 * Q;
    static constexpr int O_SHI = O_NRI + 4 * Q;
    static constexpr int O_SLO = O_SHI + (SA ? 32 * RB : 0);
    static constexpr int O_ROT = O_SLO + (SA ? 32 * RB : 0);
    static constexpr int O_RQ = O_ROT + (ROT ? 8 : 0);
    static constexpr int O_IMG = O_RQ + (ROT ? 8 * Q : 0);
    static constexpr int IMG_DW = Q * 16 * (CPLX ? 2 : 1);          // dwords per unit
    static constexpr int PIECES = O_IMG + V * IMG_DW / 4;
    static constexpr int WORDS = PIECES * 4;                        // u32 per buffer
    static constexpr int NS = (32 * RB + 63) / 64;                  // copies per coil plane
    static constexpr int NIMG = (IMG_DW + 63) / 64;                 // copies per unit image
    static constexpr int NDMA = 4 + (EDDY ? 1 : 0) + (SA ? 2 * NS : 0) + V * NIMG;   // without the chain planes
    static_assert(4 * Q <= 64 && 8 * Q <= 64, "one copy per compartment table");
};

// The images of the V units at the lane's four voxels (rows 4*lq..4*lq+3) from a stage buffer.
template <int Q, bool CPLX, bool EDDY, int V>
__device__ __forceinline__ void read_images(Vox4<Q, CPLX, EDDY, V>& v, const uint32_t* ib, int img_dw, int lq) {
#pragma unroll
    for (int vv = 0; vv < V; vv++) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            if constexpr (CPLX) {
                const f32x4 m01 = *reinterpret_cast<const f32x4*>(ib + vv * img_dw + (q * 16 + lq * 4) * 2);
                const f32x4 m23 = *reinterpret_cast<const f32x4*>(ib + vv * img_dw + (q * 16 + lq * 4) * 2 + 4);
                v.cr[vv][q][0] = m01.x; v.ci[vv][q][0] = m01.y; v.cr[vv][q][1] = m01.z; v.ci[vv][q][1] = m01.w;
                v.cr[vv][q][2] = m23.x; v.ci[vv][q][2] = m23.y; v.cr[vv][q][3] = m23.z; v.ci[vv][q][3] = m23.w;
            } else {
                const f32x4 m = *reinterpret_cast<const f32x4*>(ib + vv * img_dw + q * 16 + lq * 4);
                v.cr[vv][q][0] = m.x; v.cr[vv][q][1] = m.y; v.cr[vv][q][2] = m.z; v.cr[vv][q][3] = m.w;
            }
        }
    }
}

// ----------------------------------------------------------------------------- y-stage, MFMA
// Block = 4 waves = one (unit set, x tile of `xt` columns).  Per column x a wave accumulates
//   G[(coil, part), line] = sum_{y, part'} A[(coil,part), (y,part')] * B[(y,part'), line]
// with A the split coil sensitivities and B = E generated in registers, 16 lines per MFMA tile,
// GP tiles per pass and RB 16-row blocks (8 coils each).  A set holds V units of one slice (its
// volumes): the phase, the decay and A are evaluated once per voxel and line for all of them.
// Occupancy bound: 4 waves/SIMD (128 VGPRs) where the instantiation fits without spilling, 2 for the
// 16-accumulator tiles of many coils over 64 lines, for 32 coils and more (their staging buffers) and
// for unit sets, else 3.
template <int Q, int RB, int GP, bool CPLX, bool EDDY, int V>
constexpr int ystage_waves() {
    if constexpr (V > 1) return (V == 2 && (CPLX ? 2 * Q : Q) <= 3) ? 3 : 2;
    // 32 coils of one compartment (the 3D train): 168 VGPRs at three waves, one register spilled
    // outside the stage loop, 1-3% faster than two waves
    if (Q == 1 && RB == 4 && GP == 4 && !CPLX && !EDDY) return 3;
    return (Q <= 3 && RB == 1 && !CPLX && !EDDY) ? 4 : ((RB * GP >= 8 || RB >= 4) ? 2 : 3);
}

template <int Q, int RB, int GP, bool CPLX, bool EDDY, int V>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ystage_waves<Q, RB, GP, CPLX, EDDY, V>())))
void ystage_mfma(const DevUnit* __restrict__ units, const DevSlice* __restrict__ slices,
                 const DevGroup* __restrict__ groups, const WorkItem* __restrict__ work,
                 const int32_t* __restrict__ usets, int xt) {
    constexpr bool CHAINS = GP == 4 && !EDDY;
    constexpr bool PAIRED = RB % 2 == 0;
    constexpr bool ECHAIN = CHAINS && V == 1 && Q == 1 && !CPLX;
    using L = YStage<Q, RB, CPLX, EDDY, V, CHAINS>;
    // All of the block's LDS in one object (per-wave double buffers of the stage operands, then the
    // line table): with a single LDS variable the compiler keeps the alias scopes that let the
    // stage loop read one buffer while the copy into the other is in flight.
    struct Smem {
        uint32_t ybuf[4][2][L::WORDS];
        LineRec lines[256];
    };
    __shared__ __attribute__((aligned(16))) Smem sm;
    auto& ybuf = sm.ybuf;
    LineRec* const lines_s = sm.lines;
    const WorkItem wi = work[blockIdx.x];
    int uid[V];
    if constexpr (V == 1) {
        uid[0] = wi.unit;
    } else {
#pragma unroll
        for (int vv = 0; vv < V; vv++) uid[vv] = usets[wi.unit * V + vv];
    }
    const DevUnit& u = units[uid[0]];
    const DevGroup& gr = groups[u.group];
    const DevSlice& sl = slices[u.slice];
    const void* img[V];
    float2* gbuf[V];
#pragma unroll
    for (int vv = 0; vv < V; vv++) {
        img[vv] = units[uid[vv]].images;
        gbuf[vv] = units[uid[vv]].gbuf;
    }
    for (int i = threadIdx.x; i < gr.LyP; i += 256) lines_s[i] = gr.lines[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lq = lane >> 4, ln = lane & 15;
    const int snx = gr.snx, sny = gr.sny;
    const int snyP = (sny + 15) & ~15;
    const int snxP = (snx + 15) & ~15;
    const size_t PP = (size_t)snxP * snyP;
    const int C = gr.C, LyP = gr.LyP, G = LyP >> 4;
    const int part = ln >> 3;
    const double sys_yoff = gr.sys_yoff;
    // The wave's stages: columns x = tile*xt + wave + 4k, passes of GP line tiles, 16-row chunks.
    constexpr bool img_late = V >= 4;
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int x0w = wi.tile * xt + wv;
    const int nxw = x0w < snx ? min((xt - wv + 3) / 4, (snx - x0w + 3) / 4) : 0;
    const int npass = (G + GP - 1) / GP, nch = snyP >> 4;
    const int nstage = nxw * npass * nch;
    const bool rot = (GP == 4 && !EDDY) && gr.rot_shared;
    const uint32_t P = (uint32_t)(snx * sny);
    // lane-constant byte offsets of the copies
    const uint32_t o16 = (uint32_t)lane * 16u;                                           // contiguous pieces
    const uint32_t oq = ((uint32_t)(lane >> 2) * (uint32_t)PP + 4u * (uint32_t)(lane & 3)) * 4u;   // [q][16]
    const uint32_t oc = ((uint32_t)(lane >> 2) * (uint32_t)snyP + 4u * (uint32_t)(lane & 3)) * 4u; // [coil][16]
    const uint32_t cm = CPLX ? 2u : 1u;
    // image dword d of a unit: (q, row, part) = (d / (16 cm), (d / cm) % 16, d % cm)
    auto img_off = [&](int d, int x, int ch) -> uint32_t {
        const int q = d / (16 * (int)cm), yy = (d / (int)cm) & 15, pt = d % (int)cm;
        const int y = min(ch + yy, sny - 1);
        return (((uint32_t)q * P + (uint32_t)(x + snx * y)) * cm + (uint32_t)pt) * 4u;
    };
    auto stage_copy = [&](int x, int ch, uint32_t* buf) {
        const uint32_t xy = (uint32_t)(x * snyP + ch);
        if (lane < 8) lds_copy<16>(sl.base, xy * 8u + o16, buf + L::O_BASE * 4);
        if constexpr (EDDY) {
            if (lane < 8) lds_copy<16>(sl.poly, xy * 8u + o16, buf + L::O_POLY * 4);
        }
        if (lane < 4) lds_copy<16>(sl.rate, xy * 4u + o16, buf + L::O_RATE * 4);
        if (lane < 4 * Q) lds_copy<16>(sl.nr2, xy * 4u + oq, buf + L::O_NR2 * 4);
        if (lane < 4 * Q) lds_copy<16>(sl.nri, xy * 4u + oq, buf + L::O_NRI * 4);
        if constexpr (L::SA) {
            const uint32_t xs = (uint32_t)(x * (RB * 8) * snyP + ch) * 4u;
#pragma unroll
            for (int k = 0; k < L::NS; k++) {
                const int nk = min(64, 32 * RB - 64 * k);
                const uint32_t ok = xs + (uint32_t)(k * 16 * snyP) * 4u + oc;
                if (lane < nk) lds_copy<16>(sl.s_hi, ok, buf + (L::O_SHI + 64 * k) * 4);
                if (lane < nk) lds_copy<16>(sl.s_lo, ok, buf + (L::O_SLO + 64 * k) * 4);
            }
        }
        if constexpr (GP == 4 && !EDDY) {
            if (rot) {
                if (lane < 8) lds_copy<16>(sl.rot, xy * 8u + o16, buf + L::O_ROT * 4);
                if (lane < 8 * Q) lds_copy<16>(sl.rq, xy * 4u + oq, buf + L::O_RQ * 4);
            }
        }
#pragma unroll
        for (int vv = 0; vv < V; vv++) {
#pragma unroll
            for (int k = 0; k < L::NIMG; k++) {
                const int d = 64 * k + lane;
                if (d < L::IMG_DW) lds_copy<4>(img[vv], img_off(d, x, ch), buf + L::O_IMG * 4 + vv * L::IMG_DW + 64 * k);
            }
        }
    };
    f32x4 acc[V][RB][GP];
    int sxw = 0, spi = 0, sch = 0;            // the stage being evaluated
    if (nstage > 0) stage_copy(x0w, 0, &ybuf[wv][0][0]);
    // One stage: evaluate from buffer B while the next stage's copy lands in Bn.  The buffers and
    // the line table are restrict parameters, so the compiler knows that the reads of B and of the
    // line table do not depend on the copy in flight and waits only as far as wait_vm asks.
    auto run_stage = [&](int st, const uint32_t* __restrict__ B, uint32_t* __restrict__ Bn,
                         const LineRec* __restrict__ lines) {
        const int x = x0w + 4 * sxw, g0 = spi * GP, ch = sch * 16;
        // next stage's copy into the other buffer, then wait for this stage's
        int nxw_ = sxw, npi = spi, nchi = sch + 1;
        if (nchi == nch) { nchi = 0; if (++npi == npass) { npi = 0; ++nxw_; } }
        if (st + 1 < nstage) {
            stage_copy(x0w + 4 * nxw_, nchi * 16, Bn);
            if constexpr (GP == 4 && !EDDY) {
                if (rot) wait_vm<L::NDMA + 2>(); else wait_vm<L::NDMA>();
            } else {
                wait_vm<L::NDMA>();
            }
        } else {
            wait_vm<0>();
        }
        if (sch == 0) {
#pragma unroll
            for (int vv = 0; vv < V; vv++)
#pragma unroll
                for (int rb = 0; rb < RB; rb++)
#pragma unroll
                    for (int g = 0; g < GP; g++) acc[vv][rb][g] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
        {
                const int y0 = ch + lq * 4;
                Vox4<Q, CPLX, EDDY, V> v;
                {
                    const f64x2 b01 = *reinterpret_cast<const f64x2*>(B + L::O_BASE * 4 + lq * 8);
                    const f64x2 b23 = *reinterpret_cast<const f64x2*>(B + L::O_BASE * 4 + lq * 8 + 4);
                    v.base[0] = b01.x; v.base[1] = b01.y; v.base[2] = b23.x; v.base[3] = b23.y;
                    if constexpr (EDDY) {
                        const f64x2 p01 = *reinterpret_cast<const f64x2*>(B + L::O_POLY * 4 + lq * 8);
                        const f64x2 p23 = *reinterpret_cast<const f64x2*>(B + L::O_POLY * 4 + lq * 8 + 4);
                        v.poly[0] = p01.x; v.poly[1] = p01.y; v.poly[2] = p23.x; v.poly[3] = p23.y;
                    }
                    const f32x4 r4 = *reinterpret_cast<const f32x4*>(B + L::O_RATE * 4 + lq * 4);
                    v.rate[0] = r4.x; v.rate[1] = r4.y; v.rate[2] = r4.z; v.rate[3] = r4.w;
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        const f32x4 a = *reinterpret_cast<const f32x4*>(B + L::O_NR2 * 4 + q * 16 + lq * 4);
                        const f32x4 b = *reinterpret_cast<const f32x4*>(B + L::O_NRI * 4 + q * 16 + lq * 4);
                        v.nr2[q][0] = a.x; v.nr2[q][1] = a.y; v.nr2[q][2] = a.z; v.nr2[q][3] = a.w;
                        v.nri[q][0] = b.x; v.nri[q][1] = b.y; v.nri[q][2] = b.z; v.nri[q][3] = b.w;
                    }
                    // padded rows (y >= sny) read the clamped last row: finite E against zero coil rows.
                    // Sets of four re-read the images per line (img_late) so only one line's worth of
                    // them is live next to the four units' accumulators.
                    if constexpr (!img_late) read_images(v, B + L::O_IMG * 4, L::IMG_DW, lq);
                }
                // A operands.  16 coils and more (RB even): tile 2k holds the real rows [S_re, -S_im]
                // and tile 2k+1 the imaginary rows [S_im, S_re] of coils 16k..16k+15 (lane row =
                // coil), both formed from one read at one op per dword, and the two tiles'
                // accumulators are (re, im) of the same coils.  Up to 8 coils: one tile of rows
                // (coil, part), part = bit 3 of the lane.
                u32x4 ahi[RB], alo[RB];
#pragma unroll
                for (int rp = 0; rp < (PAIRED ? RB / 2 : RB); rp++) {
                    const int coil = PAIRED ? rp * 16 + ln : rp * 8 + (ln & 7);
                    u32x4 h, l;
                    if constexpr (L::SA) {
                        const int so = coil * 16 + lq * 4;
                        h = *reinterpret_cast<const u32x4*>(B + L::O_SHI * 4 + so);
                        l = *reinterpret_cast<const u32x4*>(B + L::O_SLO * 4 + so);
                    } else {
                        // split maps are padded to RB*8 coils with zeros: no range check
                        const uint32_t so = (uint32_t)((x * (RB * 8) + coil) * snyP + y0);
                        h = *at32<u32x4>(sl.s_hi, so);
                        l = *at32<u32x4>(sl.s_lo, so);
                    }
                    if constexpr (PAIRED) {
                        ahi[2 * rp] = rows_of(h, 0);
                        alo[2 * rp] = rows_of(l, 0);
                        ahi[2 * rp + 1] = rows_of(h, 1);
                        alo[2 * rp + 1] = rows_of(l, 1);
                    } else {
                        ahi[rp] = rows_of(h, part);
                        alo[rp] = rows_of(l, part);
                    }
                }
                const double yy0 = (double)y0 - sys_yoff;
                // Chain passes: the lane's four lines form an arithmetic progression in (t, kyn, trf,
                // tau) with the group's shared steps.  Line 0 is evaluated directly; lines 1..3
                // advance the phase by the voxel's rotor cis(2 pi (rate dt + dkyn yy)) and each decay
                // by the voxel's exp2(dtrf nr2 + dtau nri), both precomputed by the slice prep.  With
                // one compartment of real images (ECHAIN) the chain carries E itself, stepped by the
                // rotor times the decay step.
                const bool chain = CHAINS && (g0 >> 2) < gr.rotor_passes;
                f32x4 srot0, srot1, srq[Q];
                auto chain_steps = [&](int cls) {
                    srot0 = *reinterpret_cast<const f32x4*>(B + L::O_ROT * 4 + lq * 8);
                    srot1 = *reinterpret_cast<const f32x4*>(B + L::O_ROT * 4 + lq * 8 + 4);
#pragma unroll
                    for (int q = 0; q < Q; q++)
                        srq[q] = *reinterpret_cast<const f32x4*>(B + L::O_RQ * 4 + (cls * Q + q) * 16 + lq * 4);
                };
                // One loop for both kinds of pass where the accumulators are many (paired tiles) or the
                // chain state is small (ECHAIN); otherwise separate loops, which keep fewer operands live.
                if constexpr (V == 1 && (PAIRED || ECHAIN)) {
                    if (CHAINS && chain) chain_steps(lines[g0 * 16 + ln].cls);
                    float ps[4], pc[4], dq[4][Q];      // phase and decays of the current line
                    float zr[4], zi[4], sr[4], si[4];  // ECHAIN: E of the current line and its step
#pragma unroll
                    for (int g = 0; g < GP; g++) {
                        // Every tile runs the same MFMA block whichever way B was generated, so the
                        // accumulators stay in place.
                        u32x4 bh[V], bl[V];
#pragma unroll
                        for (int vv = 0; vv < V; vv++) bh[vv] = bl[vv] = u32x4{0u, 0u, 0u, 0u};
                        if (g0 + g < G) {
                            if constexpr (img_late) {
                                asm volatile("" ::: "memory");   // keep the reads in this line's iteration
                                read_images(v, B + L::O_IMG * 4, L::IMG_DW, lq);
                            }
                            if (CHAINS && g > 0 && chain) {
#pragma unroll
                                for (int j = 0; j < 4; j++) {
                                    if constexpr (ECHAIN) {
                                        const float nr = zr[j] * sr[j] - zi[j] * si[j];
                                        zi[j] = __builtin_fmaf(zr[j], si[j], zi[j] * sr[j]);
                                        zr[j] = nr;
                                    } else {
                                        const float rc = (j < 2 ? srot0 : srot1)[2 * (j & 1)];
                                        const float rs = (j < 2 ? srot0 : srot1)[2 * (j & 1) + 1];
                                        const float npc = pc[j] * rc - ps[j] * rs;
                                        ps[j] = __builtin_fmaf(pc[j], rs, ps[j] * rc);
                                        pc[j] = npc;
#pragma unroll
                                        for (int q = 0; q < Q; q++) dq[j][q] *= srq[q][j];
                                    }
#pragma unroll
                                    for (int vv = 0; vv < V; vv++) {
                                        float er, ei;
                                        if constexpr (ECHAIN) {
                                            er = zr[j];
                                            ei = zi[j];
                                        } else {
                                            unit_e(v, vv, j, ps[j], pc[j], dq[j], er, ei);
                                        }
                                        uint32_t h, l;
                                        split2(er, ei, h, l);
                                        bh[vv][j] = h;
                                        bl[vv][j] = l;
                                    }
                                }
                            } else {
                                const LineRec& Lr = lines[(g0 + g) * 16 + ln];
                                const bool ok = Lr.ky >= 0;
                                const double t_s = Lr.t_s, kyn = Lr.kyn, d = Lr.d;
                                const float trf = Lr.trf, tau = Lr.tau;
#pragma unroll
                                for (int j = 0; j < 4; j++) {
                                    phase_decay(v, j, yy0 + (double)j, t_s, kyn, d, trf, tau, ps[j], pc[j], dq[j]);
                                    if constexpr (ECHAIN) {
                                        const float w = v.cr[0][0][j] * dq[j][0];
                                        zr[j] = w * pc[j];
                                        zi[j] = w * ps[j];
                                        if constexpr (CHAINS) {
                                            if (g == 0 && chain) {
                                                const float rc = (j < 2 ? srot0 : srot1)[2 * (j & 1)];
                                                const float rs = (j < 2 ? srot0 : srot1)[2 * (j & 1) + 1];
                                                sr[j] = rc * srq[0][j];
                                                si[j] = rs * srq[0][j];
                                            }
                                        }
                                    }
#pragma unroll
                                    for (int vv = 0; vv < V; vv++) {
                                        float er, ei;
                                        if constexpr (ECHAIN) {
                                            er = zr[j];
                                            ei = zi[j];
                                        } else {
                                            unit_e(v, vv, j, ps[j], pc[j], dq[j], er, ei);
                                        }
                                        uint32_t h, l;
                                        split2(ok ? er : 0.0f, ok ? ei : 0.0f, h, l);
                                        bh[vv][j] = h;
                                        bl[vv][j] = l;
                                    }
                                }
                            }
                        }
                        // tiles past the last line (the short last pass of a group whose line tiles
                        // are no multiple of GP) skip their MFMAs: uniform, and never stored
                        if (g0 + g < G) {
#pragma unroll
                            for (int vv = 0; vv < V; vv++)
#pragma unroll
                                for (int rb = 0; rb < RB; rb++)
                                    acc[vv][rb][g] = mfma3(ahi[rb], alo[rb], bh[vv], bl[vv], acc[vv][rb][g]);
                        }
                    }
                } else {
                    if (chain) {
                        const LineRec& L0 = lines[g0 * 16 + ln];
                        chain_steps(L0.cls);   // loaded here: live only on this path
                        float ps[4], pc[4], dq[4][Q];
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            phase_decay(v, j, yy0 + (double)j, L0.t_s, L0.kyn, 0.0, L0.trf, L0.tau, ps[j], pc[j],
                                        dq[j]);
#pragma unroll
                        for (int g = 0; g < GP; g++) {
                            if constexpr (img_late) {
                                asm volatile("" ::: "memory");   // keep the reads in this line's iteration
                                read_images(v, B + L::O_IMG * 4, L::IMG_DW, lq);
                            }
                            u32x4 bh[V], bl[V];
#pragma unroll
                            for (int j = 0; j < 4; j++) {
                                if (g) {
                                    const float rc = (j < 2 ? srot0 : srot1)[2 * (j & 1)];
                                    const float rs = (j < 2 ? srot0 : srot1)[2 * (j & 1) + 1];
                                    const float npc = pc[j] * rc - ps[j] * rs;
                                    ps[j] = __builtin_fmaf(pc[j], rs, ps[j] * rc);
                                    pc[j] = npc;
#pragma unroll
                                    for (int q = 0; q < Q; q++) dq[j][q] *= srq[q][j];
                                }
#pragma unroll
                                for (int vv = 0; vv < V; vv++) {
                                    float er, ei;
                                    unit_e(v, vv, j, ps[j], pc[j], dq[j], er, ei);
                                    uint32_t h, l;
                                    split2(er, ei, h, l);
                                    bh[vv][j] = h;
                                    bl[vv][j] = l;
                                }
                            }
#pragma unroll
                            for (int vv = 0; vv < V; vv++)
#pragma unroll
                                for (int rb = 0; rb < RB; rb++)
                                    acc[vv][rb][g] = mfma3(ahi[rb], alo[rb], bh[vv], bl[vv], acc[vv][rb][g]);
                        }
                    } else {
#pragma unroll
                        for (int g = 0; g < GP; g++) {
                            if (g0 + g < G) {
                                const LineRec& Lr = lines[(g0 + g) * 16 + ln];
                                const bool ok = Lr.ky >= 0;
                                if constexpr (img_late) {
                                    asm volatile("" ::: "memory");
                                    read_images(v, B + L::O_IMG * 4, L::IMG_DW, lq);
                                }
                                u32x4 bh[V], bl[V];
#pragma unroll
                                for (int j = 0; j < 4; j++) {
                                    float sn, cs, dqj[Q];
                                    phase_decay(v, j, yy0 + (double)j, Lr.t_s, Lr.kyn, Lr.d, Lr.trf, Lr.tau, sn, cs,
                                                dqj);
#pragma unroll
                                    for (int vv = 0; vv < V; vv++) {
                                        float er, ei;
                                        unit_e(v, vv, j, sn, cs, dqj, er, ei);
                                        uint32_t h, l;
                                        split2(ok ? er : 0.0f, ok ? ei : 0.0f, h, l);
                                        bh[vv][j] = h;
                                        bl[vv][j] = l;
                                    }
                                }
#pragma unroll
                                for (int vv = 0; vv < V; vv++)
#pragma unroll
                                    for (int rb = 0; rb < RB; rb++)
                                        acc[vv][rb][g] = mfma3(ahi[rb], alo[rb], bh[vv], bl[vv], acc[vv][rb][g]);
                            }
                        }
                    }
                }
        }
        if (sch == nch - 1) {
        // acc[vv][rb][g], line 16*g + ln: lane holds rows 4*lq..4*lq+3 of the tile: paired tiles,
        // coils 16*(rb/2) + 4*lq + r (rb even: real parts, odd: imaginary parts); else part lq/2 of
        // coils rb*8 + 4*(lq&1) + r.
#pragma unroll
        for (int g = 0; g < GP; g++) {
            if (g0 + g < G) {
                const int lp = (g0 + g) * 16 + ln;
#pragma unroll
                for (int vv = 0; vv < V; vv++) {
                    // G is stored split (the x-stage's B operand): {hi, lo} per complex value at
                    // [x][line*C + coil]
                    GAS u32x2* gg = reinterpret_cast<GAS u32x2*>(gst(gbuf[vv]));
                    if constexpr (PAIRED) {
#pragma unroll
                        for (int cp = 0; cp < RB / 2; cp++) {
                            const int c0 = cp * 16 + 4 * lq;
                            const int n0 = lp * C + c0;
                            const f32x4 re = acc[vv][2 * cp][g], im = acc[vv][2 * cp + 1][g];
                            u32x2 s[4];
#pragma unroll
                            for (int r = 0; r < 4; r++) s[r] = split_g(re[r], im[r]);
                            if (c0 + 4 <= C && (C & 1) == 0) {
                                // 16-byte aligned: the buffer is, and x*LyP*C + n0 is even
                                GAS u32x4* d4 = reinterpret_cast<GAS u32x4*>(gg + (size_t)x * (LyP * C) + n0);
                                d4[0] = u32x4{s[0].x, s[0].y, s[1].x, s[1].y};
                                d4[1] = u32x4{s[2].x, s[2].y, s[3].x, s[3].y};
                            } else {
#pragma unroll
                                for (int r = 0; r < 4; r++)
                                    if (c0 + r < C) gg[(size_t)x * (LyP * C) + n0 + r] = s[r];
                            }
                        }
                    } else {
                        // Lanes lq < 2 hold the real parts of coils c0..c0+3 and the lanes 32 above
                        // them the imaginary parts: the lower lane keeps coils c0, c0+1 and the upper
                        // one c0+2, c0+3, each taking the two parts it lacks from its partner.
                        const bool lower = lq < 2;
#pragma unroll
                        for (int rb = 0; rb < RB; rb++) {
                            const f32x4 v = acc[vv][rb][g];
                            const float r0 = __shfl_xor(lower ? v[2] : v[0], 32);
                            const float r1 = __shfl_xor(lower ? v[3] : v[1], 32);
                            const int c = rb * 8 + 4 * (lq & 1) + (lower ? 0 : 2);
                            const u32x2 s0 = lower ? split_g(v[0], r0) : split_g(r0, v[2]);
                            const u32x2 s1 = lower ? split_g(v[1], r1) : split_g(r1, v[3]);
                            if (c < C) gg[(size_t)x * (LyP * C) + lp * C + c] = s0;
                            if (c + 1 < C) gg[(size_t)x * (LyP * C) + lp * C + c + 1] = s1;
                        }
                    }
                }
            }
        }
        }
        sxw = nxw_; spi = npi; sch = nchi;
    };
    // The line table goes in through an offset the compiler cannot fold (a scalar zero from asm): a
    // restrict parameter bound to the array itself would get no alias scope.
    uint32_t zero;
    asm volatile("s_mov_b32 %0, 0" : "=s"(zero));
    const LineRec* lines_p = lines_s + zero;
    for (int st = 0; st < nstage; st++) run_stage(st, &ybuf[wv][st & 1][0], &ybuf[wv][(st + 1) & 1][0], lines_p);
}

// ----------------------------------------------------------------------------- x-stage, MFMA
// Persistent blocks, as many as are resident at once, over a list of items (unit, block of 16*MT kx
// rows, share `part` of the unit's column tiles).  Block b takes the contiguous item range
// [b*nwork/grid, (b+1)*nwork/grid) of a list sorted so that neighbours share a twiddle set; it
// stages the packed twiddle rows (hi and lo; both ghost tables when the group has a ghost) and the
// group's line facts in LDS, and stages again only when the set changes.  Each wave streams
// 16-column tiles n = line*C + coil of G through them.  The complex product is carried by the B
// operand: for each 16-row kx tile, one MFMA chain against B_re = [G_re, -G_im] gives Re K and one
// against B_im = [G_im, G_re] gives Im K, so the twiddles enter as stored.
//
// Every wave walks the same sequence of steps (item, tile slot j, batch k of XB 16-x chunks): an
// item has ceil(tiles / (4 nparts)) slots for every wave, a slot past the unit's tiles loading and
// storing nothing, so the block's waves meet at the restaging barriers together.  The loads of the
// next step are in flight while the current one is multiplied, across item boundaries.  Every step
// issues all of its loads, a chunk past the last x or a step past the block's last item reading
// from an offset beyond the buffer (the load returns zeros and touches no memory): with no load
// under a branch the compiler's wait counts stay exact.  G arrives split ({hi, lo} per complex value,
// written so by the y-stage): the loads are the B operands.
// NW waves per block: 4, or 8 where the twiddle rows of more kx (up to 96) fill most of a CU's LDS
// (one block per CU, two waves per SIMD) so that fewer kx blocks re-read each unit's G.  XB: the
// 16-x chunks per step (fewer above MT = 4, where the accumulators and twiddle operands grow).
template <int MT, int NW = 4>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(2))) void xstage_mfma(const DevUnit* __restrict__ units,
                                                   const DevGroup* __restrict__ groups,
                                                   const WorkItem* __restrict__ work, int nwork, int nparts) {
    extern __shared__ uint32_t tl[];
    constexpr int rows = 16 * MT;
    constexpr int XB = MT > 4 ? 4 : 8;
    constexpr int NT = 64 * NW;   // threads per block
    constexpr uint32_t kOutside = 0x80000000u;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lq = lane >> 4, ln = lane & 15;
    const int i_lo = (int)(((int64_t)blockIdx.x * nwork) / gridDim.x);
    const int i_hi = (int)(((int64_t)(blockIdx.x + 1) * nwork) / gridDim.x);
    if (i_lo >= i_hi) return;

    // A step cursor and its item's facts (wave-uniform).  Past the last item the facts stay those of
    // the last one and the step loads nothing.
    struct Cur {
        int i, j, k;
        int group, kxb, nt0;
        int C, LyP, snxP, nx, ny, ntiles, nch, nbt, nslot;
        const float2* gbuf;
        float2* out;
    };
    auto enter = [&](Cur& c, int i) {
        c.i = i;
        c.j = 0;
        c.k = 0;
        if (i >= i_hi) return;
        const WorkItem wi = work[i];
        const DevUnit& u = units[wi.unit];
        const DevGroup& gr = groups[u.group];
        c.group = u.group;
        c.kxb = wi.tile / nparts;
        const int part = wi.tile % nparts;
        c.C = gr.C;
        c.LyP = gr.LyP;
        c.snxP = (gr.snx + 15) & ~15;
        c.nx = gr.nx;
        c.ny = gr.ny;
        c.ntiles = (gr.LyP * gr.C) >> 4;
        c.nch = c.snxP >> 4;
        c.nbt = (c.nch + XB - 1) / XB;
        c.nt0 = part * NW + wv;
        c.nslot = (c.ntiles - part * NW + NW * nparts - 1) / (NW * nparts);
        c.gbuf = u.gbuf;
        c.out = u.out;
    };
    auto advance = [&](Cur& c) {
        if (c.i >= i_hi) return;
        if (++c.k < c.nbt) return;
        c.k = 0;
        if (++c.j < c.nslot) return;
        enter(c, c.i + 1);
    };

    // Staging of the twiddle rows (SU loads of each kind in flight per thread: one latency per SU
    // rows of work rather than per row) and of the line facts: per padded line 2*ky + table
    // (negative for padding rows), read by the table choice and the stores.
    int s_group = -1, s_kxb = -1, s_ntab = 1;
    auto stage = [&](const Cur& c) {
        const DevGroup& gr = groups[c.group];
        const int snxP = c.snxP, nxP = (c.nx + 7) & ~7, stride = snxP + 4;
        const int ntab = gr.t_hi[1] != gr.t_hi[0] ? 2 : 1;
        const int kx0 = c.kxb * rows;
        constexpr int SU = 8;
        const int per_tab = rows * (snxP / 4);
        for (int tb = 0; tb < ntab; tb++) {
            for (int i0 = threadIdx.x; i0 < per_tab; i0 += NT * SU) {
                u32x4 h[SU], l[SU];
#pragma unroll
                for (int k = 0; k < SU; k++) {
                    const int i = i0 + k * NT;
                    const int r = i / (snxP / 4), x4 = (i % (snxP / 4)) * 4;
                    const int kx = kx0 + r;
                    h[k] = l[k] = u32x4{0u, 0u, 0u, 0u};
                    if (i < per_tab && kx < nxP) {
                        h[k] = *at32<u32x4>(gr.t_hi[tb], (uint32_t)(kx * snxP + x4));
                        l[k] = *at32<u32x4>(gr.t_lo[tb], (uint32_t)(kx * snxP + x4));
                    }
                }
#pragma unroll
                for (int k = 0; k < SU; k++) {
                    const int i = i0 + k * NT;
                    if (i < per_tab) {
                        const int r = i / (snxP / 4), x4 = (i % (snxP / 4)) * 4;
                        *reinterpret_cast<u32x4*>(tl + ((tb * 2 + 0) * rows + r) * stride + x4) = h[k];
                        *reinterpret_cast<u32x4*>(tl + ((tb * 2 + 1) * rows + r) * stride + x4) = l[k];
                    }
                }
            }
        }
        int32_t* const lrow = reinterpret_cast<int32_t*>(tl + ntab * 2 * rows * stride);
        for (int i = threadIdx.x; i < c.LyP; i += NT) lrow[i] = gr.lines[i].ky * 2 + (gr.lines[i].table & 1);
        s_group = c.group;
        s_kxb = c.kxb;
        s_ntab = ntab;
    };

    // The step's loads: G through a buffer resource, the row offset x*(LyP*C) (uniform) in the scalar
    // offset and the lane's (its 4-x group, column) in the vector one, so no load needs address
    // arithmetic.
    auto load_step = [&](const Cur& c, u32x2 (&buf)[XB][4]) {
        const int nt = c.nt0 + c.j * NW * nparts;
        const bool valid = c.i < i_hi && nt < c.ntiles;
        const __amdgpu_buffer_rsrc_t grs = __builtin_amdgcn_make_buffer_rsrc(
            (void*)c.gbuf, (short)0, (int)((uint32_t)c.snxP * (uint32_t)(c.LyP * c.C) * 8u), 0x00020000);
        const uint32_t xst = (uint32_t)(c.LyP * c.C);
        const uint32_t off = valid ? ((uint32_t)(lq * 4) * xst + (uint32_t)(nt * 16 + ln)) * 8u : kOutside;
#pragma unroll
        for (int cc = 0; cc < XB; cc++) {
            const int ch = c.k * XB + cc;
            const uint32_t o = ch < c.nch ? off : kOutside;
            const uint32_t row = (uint32_t)(min(ch, c.nch - 1) * 16);
#pragma unroll
            for (int jj = 0; jj < 4; jj++)
                buf[cc][jj] = __builtin_bit_cast(
                    u32x2, __builtin_amdgcn_raw_buffer_load_b64(grs, o, (row + jj) * xst * 8u, 0));
        }
    };

    f32x4 accr[MT], acci[MT];
    auto compute_step = [&](const Cur& c, const u32x2 (&buf)[XB][4]) {
        if (c.j == 0 && c.k == 0 && (c.group != s_group || c.kxb != s_kxb)) {
            // every wave is at this step: none still reads the rows being replaced
            __syncthreads();
            stage(c);
            __syncthreads();
        }
        const int stride = c.snxP + 4;
        const int nt = c.nt0 + c.j * NW * nparts;
        const bool valid = nt < c.ntiles;
        const int32_t* const lrow = reinterpret_cast<const int32_t*>(tl + s_ntab * 2 * rows * stride);
        const int tb = (s_ntab == 2 && valid) ? __builtin_amdgcn_readfirstlane(lrow[(nt * 16) / c.C] & 1) : 0;
        const uint32_t* __restrict__ th = tl + (tb * 2 + 0) * rows * stride;
        const uint32_t* __restrict__ tlo = tl + (tb * 2 + 1) * rows * stride;
        if (c.k == 0) {
#pragma unroll
            for (int m = 0; m < MT; m++) accr[m] = acci[m] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int cc = 0; cc < XB; cc++) {
            const int ch = c.k * XB + cc;
            if (ch < c.nch) {
                const int x0 = ch * 16 + lq * 4;
                u32x4 bhi, blo;
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    bhi[jj] = buf[cc][jj].x;
                    blo[jj] = buf[cc][jj].y;
                }
                // B_re = [G_re, -G_im] (negate the high halves), B_im = [G_im, G_re] (swap halves)
                const u32x4 brh = rows_of(bhi, 0), brl = rows_of(blo, 0);
                const u32x4 bih = rows_of(bhi, 1), bil = rows_of(blo, 1);
                u32x4 h[MT], l[MT];
#pragma unroll
                for (int m = 0; m < MT; m++) {
                    const int r = m * 16 + ln;
                    h[m] = *reinterpret_cast<const u32x4*>(th + r * stride + x0);
                    l[m] = *reinterpret_cast<const u32x4*>(tlo + r * stride + x0);
                }
                // the three split terms, each over all accumulators before the next (every
                // accumulator still sees lo*hi, hi*lo, hi*hi in that order)
#pragma unroll
                for (int m = 0; m < MT; m++) {
                    accr[m] = mfma_bf16(l[m], brh, accr[m]);
                    acci[m] = mfma_bf16(l[m], bih, acci[m]);
                }
#pragma unroll
                for (int m = 0; m < MT; m++) {
                    accr[m] = mfma_bf16(h[m], brl, accr[m]);
                    acci[m] = mfma_bf16(h[m], bil, acci[m]);
                }
#pragma unroll
                for (int m = 0; m < MT; m++) {
                    accr[m] = mfma_bf16(h[m], brh, accr[m]);
                    acci[m] = mfma_bf16(h[m], bih, acci[m]);
                }
            }
        }
        if (c.k == c.nbt - 1 && valid) {
            const int n = nt * 16 + ln;
            const int lp = n / c.C, coil = n % c.C;
            const int ky = lrow[lp] >> 1;
            if (ky >= 0) {
                const int nx = c.nx;
                GAS f32x2* orow = reinterpret_cast<GAS f32x2*>(gst(c.out) + ((size_t)coil * c.ny + ky) * nx);
#pragma unroll
                for (int m = 0; m < MT; m++) {
                    const int kx = c.kxb * rows + m * 16 + 4 * lq;
                    if ((nx & 1) == 0 && kx + 4 <= nx) {
                        // 16-byte aligned: the row starts at an even complex index and kx is a multiple of 4
                        GAS f32x4* o4 = reinterpret_cast<GAS f32x4*>(orow + kx);
                        o4[0] = f32x4{accr[m][0], acci[m][0], accr[m][1], acci[m][1]};
                        o4[1] = f32x4{accr[m][2], acci[m][2], accr[m][3], acci[m][3]};
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; r++)
                            if (kx + r < nx) orow[kx + r] = f32x2{accr[m][r], acci[m][r]};
                    }
                }
            }
        }
    };

    Cur lc, cc;
    enter(lc, i_lo);
    enter(cc, i_lo);
    stage(cc);
    __syncthreads();
    u32x2 bufa[XB][4], bufb[XB][4];
    load_step(lc, bufa);
    advance(lc);
    while (cc.i < i_hi) {
        load_step(lc, bufb);
        advance(lc);
        compute_step(cc, bufa);
        advance(cc);
        if (cc.i >= i_hi) break;
        load_step(lc, bufa);
        advance(lc);
        compute_step(cc, bufb);
        advance(cc);
    }
}

// ----------------------------------------------------------------------------- plain FP32
// The FP32 reference path: one launch pair per slice unit, all arithmetic in FP32 on the VALU (the
// phase formed and reduced in FP64), partial sums added in a fixed order (deterministic).

// Sum of YS partial results per output: part[k][o] for k = 0..YS-1, o < nout, added in k order by
// one thread per output (all threads of the block take outputs in turn).
template <int YS>
__device__ __forceinline__ float2 plain_sum(const float2* part, int nout, int o) {
    float2 t = part[o];
#pragma unroll 8
    for (int k = 1; k < YS; k++) {
        const float2 v = part[k * nout + o];
        t.x += v.x;
        t.y += v.y;
    }
    return t;
}

// y-stage: FP32 complex MACs over y for all coils.  Block = 256 threads = YS y-parts x (256 / YS)
// consecutive padded lines of one column x; grid = snx * ceil(LyP / (256 / YS)) blocks per unit (a
// 128^2 slice of 64 lines at YS = 32: 1024 blocks, 4 y per thread).  The y loop runs in pairs whose
// loads are all issued before either value's arithmetic.
template <int CM, int YS>
__global__ __launch_bounds__(256) void ystage_plain(const DevUnit* __restrict__ units,
                                                    const DevSlice* __restrict__ slices,
                                                    const DevGroup* __restrict__ groups, int unit) {
    constexpr int NL = 256 / YS;   // lines per block
    __shared__ float2 part[YS * NL * CM];
    const DevUnit& u = units[unit];
    const DevGroup& gr = groups[u.group];
    const DevSlice& sl = slices[u.slice];
    const int LyP = gr.LyP, snx = gr.snx, sny = gr.sny;
    const int nlb = (LyP + NL - 1) / NL;
    const int x = blockIdx.x / nlb, l0 = (blockIdx.x % nlb) * NL;
    const int li = threadIdx.x % NL, g = threadIdx.x / NL;
    const int lp = l0 + li;
    const int C = gr.C, Q = gr.Q;
    const int snyP = (sny + 15) & ~15, snxP = (snx + 15) & ~15;
    const size_t PP = (size_t)snxP * snyP, P = (size_t)snx * sny;
    const bool cplx = gr.cplx != 0;
    float2 acc[CM];
#pragma unroll
    for (int c = 0; c < CM; c++) acc[c] = make_float2(0.f, 0.f);
    const int ylen = (sny + YS - 1) / YS, y0 = g * ylen, y1 = min(sny, y0 + ylen);
    if (x < snx && lp < LyP && gr.lines[lp].ky >= 0) {
        const LineRec L = gr.lines[lp];
        // E of voxel (x, y) at this line: the compartment sum with its decays, times cis(2 pi cyc)
        auto e_of = [&](int y, float& er, float& ei) {
            const size_t o = (size_t)x * snyP + y;
            const size_t i = (size_t)x + (size_t)snx * y;
            double cyc = __builtin_fma(L.kyn, (double)y - gr.sys_yoff, sl.base[o]);
            cyc = __builtin_fma((double)sl.rate[o], L.t_s, cyc);
            cyc = __builtin_fma(sl.poly[o], L.d, cyc);
            float sn, cs;
            sincos_turns(cyc, sn, cs);
            float wr = 0.f, wi = 0.f;
            for (int q = 0; q < Q; q++) {
                const float dq =
                    __builtin_amdgcn_exp2f(__builtin_fmaf(L.trf, sl.nr2[q * PP + o], L.tau * sl.nri[q * PP + o]));
                if (cplx) {
                    const float2 m = reinterpret_cast<const float2*>(u.images)[q * P + i];
                    wr = __builtin_fmaf(m.x, dq, wr);
                    wi = __builtin_fmaf(m.y, dq, wi);
                } else {
                    wr = __builtin_fmaf(reinterpret_cast<const float*>(u.images)[q * P + i], dq, wr);
                }
            }
            er = wr * cs - wi * sn;
            ei = wr * sn + wi * cs;
        };
        auto mac = [&](int y, float er, float ei) {
#pragma unroll
            for (int c = 0; c < CM; c++) {
                if (c < C) {
                    const float2 s = sl.s32[((size_t)c * snxP + x) * snyP + y];
                    acc[c].x = __builtin_fmaf(s.x, er, __builtin_fmaf(-s.y, ei, acc[c].x));
                    acc[c].y = __builtin_fmaf(s.x, ei, __builtin_fmaf(s.y, er, acc[c].y));
                }
            }
        };
        int y = y0;
        for (; y + 1 < y1; y += 2) {
            float er0, ei0, er1, ei1;
            e_of(y, er0, ei0);
            e_of(y + 1, er1, ei1);
            mac(y, er0, ei0);
            mac(y + 1, er1, ei1);
        }
        if (y < y1) {
            float er, ei;
            e_of(y, er, ei);
            mac(y, er, ei);
        }
    }
#pragma unroll
    for (int c = 0; c < CM; c++) part[(g * NL + li) * CM + c] = acc[c];
    __syncthreads();
    // outputs (line li, coil c) = part index li*CM + c; the YS parts of one output lie NL*CM apart
    for (int o = threadIdx.x; o < NL * CM; o += 256) {
        const int l = o / CM, c = o % CM;
        if (c < C && x < snx && l0 + l < LyP) u.gbuf[((size_t)x * LyP + l0 + l) * C + c] = plain_sum<YS>(part, NL * CM, o);
    }
}

// x-stage: FP32 on the VALU.  Block = 256 threads = XS x-parts x (256 / XS) consecutive (column n =
// line*C + coil, kx) outputs; grid = ceil(LyP*C*nx / (256 / XS)) blocks (a 64-line 8-coil 64^2
// k-space at XS = 16: 2048 blocks, 8 x per thread).
template <int XS>
__global__ __launch_bounds__(256) void xstage_plain(const DevUnit* __restrict__ units,
                                                    const DevGroup* __restrict__ groups, int unit) {
    constexpr int NO = 256 / XS;
    __shared__ float2 part[XS * NO];
    const DevUnit& u = units[unit];
    const DevGroup& gr = groups[u.group];
    const int nx = gr.nx, C = gr.C, LyP = gr.LyP, snx = gr.snx, snxP = (snx + 15) & ~15;
    const int oi = threadIdx.x % NO, g = threadIdx.x / NO;
    const int t = blockIdx.x * NO + oi;
    const int kx = t % nx, n = t / nx;
    const bool live = n < LyP * C;
    const int lp = live ? n / C : 0, c = n % C;
    const LineRec L = gr.lines[lp];
    float ar = 0.f, ai = 0.f;
    const int xlen = (snx + XS - 1) / XS, x0 = g * xlen, x1 = min(snx, x0 + xlen);
    if (live && L.ky >= 0) {
        const float2* __restrict__ tw = gr.t32[L.table] + (size_t)kx * snxP;
        const float2* __restrict__ gcol = u.gbuf + (size_t)lp * C + c;
        const size_t gstep = (size_t)LyP * C;
#pragma unroll 4
        for (int x = x0; x < x1; x++) {
            const float2 gv = gcol[(size_t)x * gstep];
            const float2 w = tw[x];
            ar = __builtin_fmaf(gv.x, w.x, __builtin_fmaf(-gv.y, w.y, ar));
            ai = __builtin_fmaf(gv.x, w.y, __builtin_fmaf(gv.y, w.x, ai));
        }
    }
    part[g * NO + oi] = make_float2(ar, ai);
    __syncthreads();
    if (g == 0 && live && L.ky >= 0) u.out[((size_t)c * gr.ny + L.ky) * nx + kx] = plain_sum<XS>(part, NO, oi);
}

// ----------------------------------------------------------------------------- adjoint (VJP)
// Cotangent g = dL/dK (PyTorch's conjugate-Wirtinger convention), per unit.
// x-adjoint: Gb_c(x, l) = sum_kx conj(T_tab(l)[kx, x]) g_c[kx, ky_l]  (FP32).  Block = 4 waves; a wave
// owns one line and a group of XA_CG coils, its 64 lanes consecutive x.  The twiddle column reads
// are coalesced along x and reused for the group's coils; the cotangent values are uniform over the
// wave (scalar loads).  grid = (ceil(snx / 64), LyP/4 * ceil(C / XA_CG), units).
constexpr int XA_CG = 8;
__global__ __launch_bounds__(256) void xadjoint(const DevUnit* __restrict__ units, const DevGroup* __restrict__ groups,
                                                const VjpUnit* __restrict__ vj, int unit0) {
    const int unit = unit0 + blockIdx.z;
    const DevUnit& u = units[unit];
    const VjpUnit& v = vj[unit];
    const DevGroup& gr = groups[u.group];
    const int C = gr.C, LyP = gr.LyP, snx = gr.snx, snxP = (snx + 15) & ~15, nx = gr.nx, ny = gr.ny;
    const int ncg = (C + XA_CG - 1) / XA_CG;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int cg = blockIdx.y % ncg, lp = (blockIdx.y / ncg) * 4 + wave;   // LyP is a multiple of 16
    const int x = blockIdx.x * 64 + lane, c0 = cg * XA_CG;
    const LineRec L = gr.lines[lp];
    float ar[XA_CG], ai[XA_CG];
#pragma unroll
    for (int k = 0; k < XA_CG; k++) ar[k] = ai[k] = 0.0f;
    if (L.ky >= 0 && v.gout) {
        // Lanes past the row read column snx - 1 and coils past C read coil C - 1 (both in bounds);
        // neither is stored.  Four kx per step, every load of the step issued before its products.
        const GAS f32x2* tw = reinterpret_cast<const GAS f32x2*>(gld(gr.t32[L.table])) + min(x, snx - 1);
        const GAS f32x2* g = reinterpret_cast<const GAS f32x2*>(gld(v.gout)) + (size_t)L.ky * nx;
        const size_t cstride = (size_t)ny * nx;
        size_t coff[XA_CG];
#pragma unroll
        for (int k = 0; k < XA_CG; k++) coff[k] = (size_t)min(c0 + k, C - 1) * cstride;
        auto step = [&](int kx, int nk) {
            f32x2 w[4], gv[XA_CG][4];
#pragma unroll
            for (int j = 0; j < 4; j++) w[j] = j < nk ? tw[(size_t)(kx + j) * snxP] : f32x2{0.f, 0.f};
#pragma unroll
            for (int k = 0; k < XA_CG; k++)
#pragma unroll
                for (int j = 0; j < 4; j++) gv[k][j] = j < nk ? g[coff[k] + kx + j] : f32x2{0.f, 0.f};
#pragma unroll
            for (int k = 0; k < XA_CG; k++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    ar[k] = __builtin_fmaf(w[j].x, gv[k][j].x, __builtin_fmaf(w[j].y, gv[k][j].y, ar[k]));
                    ai[k] = __builtin_fmaf(w[j].x, gv[k][j].y, __builtin_fmaf(-w[j].y, gv[k][j].x, ai[k]));
                }
        };
        int kx = 0;
        for (; kx + 4 <= nx; kx += 4) step(kx, 4);
        if (kx < nx) step(kx, nx - kx);
    }
    if (x < snx) {
        float2* __restrict__ o = v.gbar + ((size_t)x * LyP + lp) * C + c0;
#pragma unroll
        for (int k = 0; k < XA_CG; k++)
            if (c0 + k < C) o[k] = make_float2(ar[k], ai[k]);
    }
}

// x-adjoint on MFMA (split-BF16 plans): Gb[x, n] = sum_kx conj(T_tab[kx, x]) g[kx, n], n = line*C + coil,
// the forward x-stage with the roles of x and kx exchanged.  A block stages the conjugated, transposed
// twiddle rows of its 16*MT x (split hi and lo, both ghost tables when the group has a ghost) in LDS
// from the FP32 tables; each wave streams its 16-column tiles of the cotangent, 16 kx per step (a lane
// reads 4 consecutive kx of one column and splits them), and the complex product rides on B as in the
// forward: B_re = [g_re, -g_im] gives Re Gb, B_im = [g_im, g_re] gives Im Gb.  Padded lines (ky < 0)
// read no cotangent and write zero rows.
// Work w = ((unit - unit0) * nparts + part) * nxb + xb, w < nwork: the nxb x blocks of one (unit, part)
// read the same cotangent columns, so they take consecutive dispatch slots of one XCD (blocks go
// round-robin over the 8 XCDs): block b runs w = (b % 8) * (gridDim.x / 8) + b / 8, and the grid is a
// multiple of 8 blocks.
template <int MT>
__global__ __launch_bounds__(256) void xadjoint_mfma(const DevUnit* __restrict__ units,
                                                     const DevGroup* __restrict__ groups,
                                                     const VjpUnit* __restrict__ vj, int unit0, int nparts, int nxb,
                                                     int nwork) {
    extern __shared__ uint32_t tla[];
    const int w = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
    if (w >= nwork) return;
    const int xb = w % nxb, part = (w / nxb) % nparts;
    const int unit = unit0 + w / (nxb * nparts);
    const DevUnit& u = units[unit];
    const VjpUnit& v = vj[unit];
    const DevGroup& gr = groups[u.group];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lq = lane >> 4, ln = lane & 15;
    const int C = gr.C, LyP = gr.LyP, snx = gr.snx, snxP = (snx + 15) & ~15, nx = gr.nx, ny = gr.ny;
    const int nkP = (nx + 15) & ~15;              // kx padded to whole 16-kx steps
    const int stride = nkP + 4;                   // u32 per LDS row (16-byte pad against bank conflicts)
    constexpr int rows = 16 * MT;
    const int x0 = xb * rows;
    const int ntab = gr.t_hi[1] != gr.t_hi[0] ? 2 : 1;
    // stage A[x][kx] = conj(T[kx][x]), split; reads run along x (coalesced), rows past snx or kx past
    // nx are zero
    const int per_tab = rows * nkP;
    for (int tb = 0; tb < ntab; tb++) {
        const float2* __restrict__ t32 = gr.t32[tb];
        for (int i = threadIdx.x; i < per_tab; i += 256) {
            const int r = i % rows, kx = i / rows;
            const int x = x0 + r;
            uint32_t h = 0u, l = 0u;
            if (x < snx && kx < nx) {
                const float2 t = t32[(size_t)kx * snxP + x];
                split2(t.x, -t.y, h, l);
            }
            tla[((tb * 2 + 0) * rows + r) * stride + kx] = h;
            tla[((tb * 2 + 1) * rows + r) * stride + kx] = l;
        }
    }
    __syncthreads();
    const int ntiles = (LyP * C) >> 4;
    const int first = part * 4 + wave, step = nparts * 4;
    const size_t cstride = (size_t)ny * nx;
    for (int nt = first; nt < ntiles; nt += step) {
        const int n = nt * 16 + ln;
        const int lp = n / C, c = n % C;
        const LineRec& L = gr.lines[lp];
        const int tb = ntab == 2 ? gr.lines[(nt * 16) / C].table : 0;   // a tile never spans two tables
        const uint32_t* __restrict__ th = tla + (tb * 2 + 0) * rows * stride;
        const uint32_t* __restrict__ tlo = tla + (tb * 2 + 1) * rows * stride;
        const bool ok = L.ky >= 0 && v.gout;
        const f32x2* __restrict__ g = reinterpret_cast<const f32x2*>(v.gout) + (ok ? (size_t)c * cstride + (size_t)L.ky * nx : 0);
        f32x4 accr[MT], acci[MT];
#pragma unroll
        for (int m = 0; m < MT; m++) accr[m] = acci[m] = f32x4{0.f, 0.f, 0.f, 0.f};
        // batches of KB 16-kx steps: the batch's cotangent loads are all in flight before the first
        // is split, one memory latency per batch rather than per step
        constexpr int KB = 4;
        for (int kb = 0; kb < nkP; kb += 16 * KB) {
            f32x2 gv[KB][4];
#pragma unroll
            for (int t = 0; t < KB; t++)
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    const int kx = kb + 16 * t + lq * 4 + jj;
                    gv[t][jj] = (ok && kx < nx) ? g[kx] : f32x2{0.f, 0.f};
                }
#pragma unroll
            for (int t = 0; t < KB; t++) {
                const int k0 = kb + 16 * t;
                if (k0 >= nkP) break;
                const int kx = k0 + lq * 4;
                u32x4 bhi, blo;
#pragma unroll
                for (int jj = 0; jj < 4; jj++) {
                    uint32_t h, l;
                    split2(gv[t][jj].x, gv[t][jj].y, h, l);
                    bhi[jj] = h;
                    blo[jj] = l;
                }
                const u32x4 brh = rows_of(bhi, 0), brl = rows_of(blo, 0);
                const u32x4 bih = rows_of(bhi, 1), bil = rows_of(blo, 1);
#pragma unroll
                for (int m = 0; m < MT; m++) {
                    const int r = m * 16 + ln;
                    const u32x4 h = *reinterpret_cast<const u32x4*>(th + r * stride + kx);
                    const u32x4 l = *reinterpret_cast<const u32x4*>(tlo + r * stride + kx);
                    accr[m] = mfma3(h, l, brh, brl, accr[m]);
                    acci[m] = mfma3(h, l, bih, bil, acci[m]);
                }
            }
        }
        // accr[m][r]: row x0 + 16 m + 4 lq + r, column n
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int x = x0 + m * 16 + 4 * lq + r;
                if (x < snx) v.gbar[((size_t)x * LyP + lp) * C + c] = make_float2(accr[m][r], acci[m][r]);
            }
    }
}

// y-adjoint: one lane per voxel (x, y), looping over the lines; a block owns 256 y of one column x,
// and each 64-line pass of the column's Gb (every lane reads the same values) is staged in LDS.
// grid = (snx * ceil(sny / 256), units).  With Eb = sum_c conj(S_c) Gb_c,
// t_l = conj(P_l) Eb_l and D_q,l = exp(-trf/T2_q - tau/T2'_q):
//   dL/dm_q   = sum_l D_q,l t_l                      (complex images; real images take the real part)
//   dL/dfmap  = 2 pi Im( sum_q conj(m_q) sum_l tb0_l D_q,l t_l )
//   dL/dT2'_q = Re( conj(m_q) sum_l tau_l D_q,l t_l ) / T2'_q^2
// Lines of a chain pass (four-line arithmetic progressions) advance P and D by rotors.
template <int Q, bool CPLX, int CM>
__global__ __launch_bounds__(256) void yadjoint(const DevUnit* __restrict__ units, const DevSlice* __restrict__ slices,
                                                const DevGroup* __restrict__ groups, const VjpUnit* __restrict__ vj,
                                                int unit0) {
    __shared__ LineRec lines_s[256];
    __shared__ float2 gs[64 * CM];
    const int unit = unit0 + blockIdx.y;
    const DevUnit& u = units[unit];
    const VjpUnit& vv = vj[unit];
    const DevGroup& gr = groups[u.group];
    const DevSlice& sl = slices[u.slice];
    for (int i = threadIdx.x; i < gr.LyP; i += 256) lines_s[i] = gr.lines[i];
    const int snx = gr.snx, sny = gr.sny, snxP = (snx + 15) & ~15, snyP = (sny + 15) & ~15;
    const int C = gr.C, LyP = gr.LyP;
    const size_t PP = (size_t)snxP * snyP, P = (size_t)snx * sny;
    const int ycb = (sny + 255) / 256;
    const int x = blockIdx.x / ycb, y0 = (blockIdx.x % ycb) * 256 + threadIdx.x;
    // lanes past the column evaluate its last voxel (in bounds), take part in the staging, store nothing
    const bool act = y0 < sny;
    const int y = act ? y0 : sny - 1;
    const size_t o = (size_t)x * snyP + y, i = (size_t)x + (size_t)snx * y;
    const double base = sl.base[o], rate = (double)sl.rate[o], poly = sl.poly[o];
    const double yy = (double)y - gr.sys_yoff;
    float nr2[Q], nri[Q], mr[Q], mi[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        nr2[q] = sl.nr2[q * PP + o];
        nri[q] = sl.nri[q * PP + o];
        if constexpr (CPLX) {
            const float2 m = reinterpret_cast<const float2*>(u.images)[q * P + i];
            mr[q] = m.x;
            mi[q] = m.y;
        } else {
            mr[q] = reinterpret_cast<const float*>(u.images)[q * P + i];
            mi[q] = 0.0f;
        }
    }
    float2 s[CM];
#pragma unroll
    for (int c = 0; c < CM; c++) s[c] = c < C ? sl.s32[((size_t)c * snxP + x) * snyP + y] : make_float2(0.f, 0.f);
    float ar[Q], ai[Q], br[Q], bi[Q], fr[Q], fi[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) ar[q] = ai[q] = br[q] = bi[q] = fr[q] = fi[q] = 0.0f;
    const float2* __restrict__ gb = vv.gbar + (size_t)x * LyP * C;
    const int G = LyP >> 4;
    for (int g0 = 0; g0 < G; g0 += 4) {
        // stage this pass's lines (up to 64) x C of the column: gs[line - 16 g0][coil]
        const int nl = min(64, LyP - g0 * 16);
        __syncthreads();   // the previous pass's reads (and, first, the line table) are done
        for (int k = threadIdx.x; k < nl * C; k += blockDim.x) gs[(k / C) * CM + k % C] = gb[(size_t)g0 * 16 * C + k];
        __syncthreads();
        const bool chain = (g0 >> 2) < gr.rotor_passes && g0 + 4 <= G;
        for (int ln = 0; ln < 16; ln++) {
            float pc = 0.f, ps = 0.f, rc = 1.f, rs = 0.f;
            float dq[Q], rq[Q];
            const int ng = chain ? 4 : (G - g0 < 4 ? G - g0 : 4);
            for (int e = 0; e < ng; e++) {
                const int lp = (g0 + e) * 16 + ln;
                const LineRec& L = lines_s[lp];
                if (L.ky < 0) continue;
                if (!chain || e == 0) {
                    double cyc = __builtin_fma(L.kyn, yy, base);
                    cyc = __builtin_fma(rate, L.t_s, cyc);
                    cyc = __builtin_fma(poly, L.d, cyc);
                    sincos_turns(cyc, ps, pc);
#pragma unroll
                    for (int q = 0; q < Q; q++)
                        dq[q] = __builtin_amdgcn_exp2f(__builtin_fmaf(L.trf, nr2[q], L.tau * nri[q]));
                    if (chain) {
                        const LineRec& L1 = lines_s[lp + 16];
                        const double dc = __builtin_fma(L1.kyn - L.kyn, yy, rate * (L1.t_s - L.t_s));
                        sincos_turns(dc, rs, rc);
#pragma unroll
                        for (int q = 0; q < Q; q++)
                            rq[q] = __builtin_amdgcn_exp2f(__builtin_fmaf(L1.trf - L.trf, nr2[q], (L1.tau - L.tau) * nri[q]));
                    }
                } else {
                    const float npc = pc * rc - ps * rs;
                    ps = __builtin_fmaf(pc, rs, ps * rc);
                    pc = npc;
#pragma unroll
                    for (int q = 0; q < Q; q++) dq[q] *= rq[q];
                }
                // Eb = sum_c conj(S_c) Gb_c
                const float2* gl = gs + (size_t)(lp - g0 * 16) * CM;
                float er = 0.f, ei = 0.f;
#pragma unroll
                for (int c = 0; c < CM; c++) {
                    if (c < C) {
                        const float2 gv = gl[c];
                        er = __builtin_fmaf(s[c].x, gv.x, __builtin_fmaf(s[c].y, gv.y, er));
                        ei = __builtin_fmaf(s[c].x, gv.y, __builtin_fmaf(-s[c].y, gv.x, ei));
                    }
                }
                // t = conj(P) Eb
                const float tr = pc * er + ps * ei, ti = pc * ei - ps * er;
                const float tb0 = L.tb0, tau = L.tau;
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const float d = dq[q];
                    ar[q] = __builtin_fmaf(d, tr, ar[q]);
                    ai[q] = __builtin_fmaf(d, ti, ai[q]);
                    const float dt = d * tau, df = d * tb0;
                    br[q] = __builtin_fmaf(dt, tr, br[q]);
                    bi[q] = __builtin_fmaf(dt, ti, bi[q]);
                    fr[q] = __builtin_fmaf(df, tr, fr[q]);
                    fi[q] = __builtin_fmaf(df, ti, fi[q]);
                }
            }
        }
    }
    if (!act) return;
    if (vv.gimg) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            if constexpr (CPLX) reinterpret_cast<float2*>(vv.gimg)[q * P + i] = make_float2(ar[q], ai[q]);
            else reinterpret_cast<float*>(vv.gimg)[q * P + i] = ar[q];
        }
    }
    if (vv.gfmap) {
        float acc = 0.f;
        // Im(conj(m) f) = mr fi - mi fr
#pragma unroll
        for (int q = 0; q < Q; q++) acc += mr[q] * fi[q] - mi[q] * fr[q];
        vv.gfmap[i] = gr.do_dist ? (float)(kTau * (double)acc) : 0.0f;
    }
    if (vv.gti) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            float ti_ms = sl.t_inhom ? sl.t_inhom[q * P + i] : (sl.has_t_inhom ? sl.ti_uniform[q] : (float)gr.t_inhom);
            const float re = mr[q] * br[q] + mi[q] * bi[q];  // Re(conj(m) b)
            vv.gti[q * P + i] = gr.do_relax ? re / (ti_ms * ti_ms) : 0.0f;
        }
    }
}

// y-adjoint on MFMA (split-BF16 plans).  The coil contraction of a column x,
//   H(y, l) = sum_c conj(S_c(x, y)) Gb_c(x, l),
// runs as split-BF16 products: A = the coil rows of 16 y (tile pair: real rows [S_re, S_im] and
// imaginary rows [-S_im, S_re] per coil), B = Gb of 16 lines (k = coil, part).  Each lane then holds
// H for its 4 voxels (rows 4 lq..4 lq + 3) at its line (column ln), evaluates their phase and decays
// there and adds the same sums as the VALU y-adjoint; the sums over the 16 lines of a tile sit in 16
// lanes and are added across them at the end.  Block = 4 waves on one (column, unit), a wave owning
// the y tiles w, w + 4, ...; NCS 16-coil steps cover C <= 16 NCS.  grid = (snx, units).
// The block first splits the column's Gb (LyP lines x C coils) once into LDS, hi and lo planes of
// rows of 16 NCS + 4 words (zero coils past C), so the waves read their B operands from LDS; the
// contraction of line tile g + 1 is issued before the VALU work of tile g.
// EDDY: the group has an eddy-current phase (the poly plane); without it that term is zero.
// Dynamic LDS: yadjoint_mfma_lds(LyP, NCS) bytes.
__host__ __device__ constexpr size_t yadjoint_mfma_lds(int LyP, int ncs) {
    return (size_t)LyP * sizeof(LineRec) + 2 * (size_t)LyP * (16 * ncs + 4) * 4;
}

// Real single-compartment images with up to 32 coils fit 128 VGPRs: four waves per SIMD.
constexpr int yadjoint_mfma_waves(int Q, bool CPLX, int NCS) { return (Q == 1 && !CPLX && NCS <= 2) ? 4 : 1; }

template <int Q, bool CPLX, int NCS, bool EDDY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(yadjoint_mfma_waves(Q, CPLX, NCS)))) void yadjoint_mfma(const DevUnit* __restrict__ units,
                                                     const DevSlice* __restrict__ slices,
                                                     const DevGroup* __restrict__ groups,
                                                     const VjpUnit* __restrict__ vj, int unit0) {
    constexpr int CW = NCS * 16, RS = CW + 4;
    extern __shared__ __align__(16) unsigned char yadj_smem[];
    const int unit = unit0 + blockIdx.y;
    const DevUnit& u = units[unit];
    const VjpUnit& vv = vj[unit];
    const DevGroup& gr = groups[u.group];
    const DevSlice& sl = slices[u.slice];
    const int snx = gr.snx, sny = gr.sny, snxP = (snx + 15) & ~15, snyP = (sny + 15) & ~15;
    const int C = gr.C, LyP = gr.LyP, G = LyP >> 4, Cp = gr.RB * 8;
    const size_t PP = (size_t)snxP * snyP, P = (size_t)snx * sny;
    const int x = blockIdx.x;
    LineRec* lines_s = reinterpret_cast<LineRec*>(yadj_smem);
    uint32_t* bsh = reinterpret_cast<uint32_t*>(yadj_smem + (size_t)LyP * sizeof(LineRec));
    uint32_t* bsl = bsh + (size_t)LyP * RS;
    const float2* __restrict__ gb = vv.gbar + (size_t)x * LyP * C;
    for (int i = threadIdx.x; i < LyP; i += 256) lines_s[i] = gr.lines[i];
    for (int e = threadIdx.x; e < LyP * CW; e += 256) {
        const int l = e / CW, c = e % CW;
        const float2 v = c < C ? gb[(size_t)l * C + c] : make_float2(0.f, 0.f);
        uint32_t h, lo;
        split2(v.x, v.y, h, lo);
        bsh[l * RS + c] = h;
        bsl[l * RS + c] = lo;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lq = lane >> 4, ln = lane & 15;
    const int nyt = (sny + 15) >> 4;
    for (int yt = wave; yt < nyt; yt += 4) {
        // A rows y = 16 yt + ln (< snyP; the split planes hold zero coils past C and zero rows past sny)
        u32x4 arh[NCS], arl[NCS];
        const int ya = yt * 16 + ln;
#pragma unroll
        for (int s = 0; s < NCS; s++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int c = s * 16 + 4 * lq + j;
                uint32_t h = 0u, l = 0u;
                if (c < Cp) {
                    const size_t so = ((size_t)x * Cp + c) * snyP + ya;
                    h = sl.s_hi[so];
                    l = sl.s_lo[so];
                }
                arh[s][j] = h;
                arl[s][j] = l;
            }
        // the lane's voxels y0 + j (rows of its accumulator tile; the planes are padded to snyP)
        const int y0 = yt * 16 + 4 * lq;
        double base[4], poly[4];
        float rate[4], nr2[Q][4], nri[Q][4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const size_t o = (size_t)x * snyP + y0 + j;
            base[j] = sl.base[o];
            poly[j] = EDDY ? sl.poly[o] : 0.0;
            rate[j] = sl.rate[o];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                nr2[q][j] = sl.nr2[q * PP + o];
                nri[q][j] = sl.nri[q * PP + o];
            }
        }
        const double yy0 = (double)y0 - gr.sys_yoff;
        float ar[Q][4], ai[Q][4], br[Q][4], bi[Q][4], fr[Q][4], fi[Q][4];
#pragma unroll
        for (int q = 0; q < Q; q++)
#pragma unroll
            for (int j = 0; j < 4; j++) ar[q][j] = ai[q][j] = br[q][j] = bi[q][j] = fr[q][j] = fi[q][j] = 0.0f;
        // H of line tile g: B = the split Gb rows of lines 16 g .. 16 g + 15 (lane column ln)
        auto contract = [&](int g, f32x4& hr, f32x4& hi) {
            const int o = (g * 16 + ln) * RS + 4 * lq;
            hr = f32x4{0.f, 0.f, 0.f, 0.f};
            hi = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < NCS; s++) {
                const u32x4 bh = *reinterpret_cast<const u32x4*>(bsh + o + 16 * s);
                const u32x4 bl = *reinterpret_cast<const u32x4*>(bsl + o + 16 * s);
                // imaginary rows [-S_im, S_re], made here rather than held in registers
                u32x4 aih, ail;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    aih[j] = __builtin_rotateright32(arh[s][j], 16) ^ 0x8000u;
                    ail[j] = __builtin_rotateright32(arl[s][j], 16) ^ 0x8000u;
                }
                hr = mfma3(arh[s], arl[s], bh, bl, hr);
                hi = mfma3(aih, ail, bh, bl, hi);
            }
        };
        for (int g = 0; g < G; g++) {
            const int lp = g * 16 + ln;
            f32x4 hr, hi;
            contract(g, hr, hi);
            const LineRec& L = lines_s[lp];
            if (L.ky >= 0) {
                const float trf = L.trf, tau = L.tau, tb0 = L.tb0;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    double cyc = __builtin_fma(L.kyn, yy0 + (double)j, base[j]);
                    cyc = __builtin_fma((double)rate[j], L.t_s, cyc);
                    if constexpr (EDDY) cyc = __builtin_fma(poly[j], L.d, cyc);
                    float ps, pc;
                    sincos_turns(cyc, ps, pc);
                    // t = conj(P) H
                    const float tr = pc * hr[j] + ps * hi[j], ti = pc * hi[j] - ps * hr[j];
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        const float d = __builtin_amdgcn_exp2f(__builtin_fmaf(trf, nr2[q][j], tau * nri[q][j]));
                        const float dt = d * tau, df = d * tb0;
                        // real images take Re of the image and T2' sums and Im of the fieldmap sum only
                        ar[q][j] = __builtin_fmaf(d, tr, ar[q][j]);
                        br[q][j] = __builtin_fmaf(dt, tr, br[q][j]);
                        fi[q][j] = __builtin_fmaf(df, ti, fi[q][j]);
                        if constexpr (CPLX) {
                            ai[q][j] = __builtin_fmaf(d, ti, ai[q][j]);
                            bi[q][j] = __builtin_fmaf(dt, ti, bi[q][j]);
                            fr[q][j] = __builtin_fmaf(df, tr, fr[q][j]);
                        }
                    }
                }
            }
        }
        // sums over the tile's 16 lines (lanes ln of one row group), then lane ln = j writes voxel y0 + j
        auto rsum = [](float v) {
#pragma unroll
            for (int m = 1; m < 16; m <<= 1) v += __shfl_xor(v, m, 16);
            return v;
        };
        float oar[Q], oai[Q], obr[Q], obi[Q], ofr[Q], ofi[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            oar[q] = oai[q] = obr[q] = obi[q] = ofr[q] = ofi[q] = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float a = rsum(ar[q][j]), b = rsum(ai[q][j]), c = rsum(br[q][j]), d = rsum(bi[q][j]);
                const float e = rsum(fr[q][j]), f = rsum(fi[q][j]);
                if (ln == j) { oar[q] = a; oai[q] = b; obr[q] = c; obi[q] = d; ofr[q] = e; ofi[q] = f; }
            }
        }
        const int y = y0 + ln;
        if (ln < 4 && y < sny) {
            const size_t i = (size_t)x + (size_t)snx * y;
            float m_r[Q], m_i[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                if constexpr (CPLX) {
                    const float2 m = reinterpret_cast<const float2*>(u.images)[q * P + i];
                    m_r[q] = m.x;
                    m_i[q] = m.y;
                } else {
                    m_r[q] = reinterpret_cast<const float*>(u.images)[q * P + i];
                    m_i[q] = 0.0f;
                }
            }
            if (vv.gimg) {
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    if constexpr (CPLX) reinterpret_cast<float2*>(vv.gimg)[q * P + i] = make_float2(oar[q], oai[q]);
                    else reinterpret_cast<float*>(vv.gimg)[q * P + i] = oar[q];
                }
            }
            if (vv.gfmap) {
                float acc = 0.f;
#pragma unroll
                for (int q = 0; q < Q; q++) acc += CPLX ? m_r[q] * ofi[q] - m_i[q] * ofr[q] : m_r[q] * ofi[q];
                vv.gfmap[i] = gr.do_dist ? (float)(kTau * (double)acc) : 0.0f;
            }
            if (vv.gti) {
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const float ti_ms = sl.t_inhom ? sl.t_inhom[q * P + i]
                                                   : (sl.has_t_inhom ? sl.ti_uniform[q] : (float)gr.t_inhom);
                    const float re = CPLX ? m_r[q] * obr[q] + m_i[q] * obi[q] : m_r[q] * obr[q];
                    vv.gti[q * P + i] = gr.do_relax ? re / (ti_ms * ti_ms) : 0.0f;
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------- z-stage (3D partitions)
// Centred z-DFT across partitions with per-excitation weights, kspace3d.rs:90-100, 677-692, 716-722:
//   forward  out[p][n] (+)= sum_c w(p, sy(n), c) sum_z zk[p][z] in_c[z][n]
//   adjoint  out_c[z][n] (+)= sum_p conj(zk[p][z]) w(p, sy(n), c) in[p][n]
// zk[p][z] = e^{+i 2 pi (p - nz/2)(z - nz/2)/nz}, symmetric in (p, z); column n = (coil, ky, kx),
// sy(n) = ky % ky_segments, w(p, sy, c) = weights[(p * ky_segments + sy) * n_comp + c].
// A block of 256 threads takes one set and ZCOLS columns; the table (split hi/lo, zero-padded to
// MT*16 square) is built once per block in LDS from an exactly reduced integer angle; each wave
// contracts 16-column tiles with the complex structure on the B operand, as the x-stage does.
static constexpr int kZCols = 256;

template <int MT, bool ADJ>
__global__ __launch_bounds__(256) void zstage(const ZSet* __restrict__ sets, const ZBlock* __restrict__ blocks) {
    constexpr int NP = MT * 16;
    __shared__ uint32_t zh[NP][NP + 4], zl[NP][NP + 4];
    __shared__ uint32_t eh[NP], el[NP];   // split e^{+-i 2 pi m / nz}, m < nz
    const ZBlock blk = blocks[blockIdx.x];
    const ZSet S = sets[blk.set];
    const int nz = S.nz, h = nz / 2;
    // the table takes nz distinct values (the angle of entry (r, k) is 2 pi m / nz exactly, with
    // m = (r - h)(k - h) mod nz): evaluate those once, then fill the table from them
    if ((int)threadIdx.x < nz) {
        double sv, cv;
        sincos(kTau * (double)threadIdx.x / (double)nz, &sv, &cv);
        uint32_t hi, lo;
        split2((float)cv, ADJ ? -(float)sv : (float)sv, hi, lo);
        eh[threadIdx.x] = hi;
        el[threadIdx.x] = lo;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < NP * NP; i += 256) {
        const int r = i / NP, k = i % NP;
        uint32_t hi = 0u, lo = 0u;
        if (r < nz && k < nz) {
            int m = ((r - h) * (k - h)) % nz;
            if (m < 0) m += nz;
            hi = eh[m];
            lo = el[m];
        }
        zh[r][k] = hi;
        zl[r][k] = lo;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lq = lane >> 4, ln = lane & 15;
    const int ncols = S.ncols;
    for (int it = 0; it < kZCols / 64; it++) {
        const int n0 = blk.col0 + (it * 4 + wave) * 16;
        if (n0 >= ncols) break;
        const int n = n0 + ln;
        const bool col_ok = n < ncols;
        const int nc = col_ok ? n : ncols - 1;
        const int sy = ((nc / S.nx) % S.ny) % S.kseg;
        if (!ADJ) {
            f32x4 outr[MT], outi[MT];
#pragma unroll
            for (int m = 0; m < MT; m++) outr[m] = outi[m] = f32x4{0.f, 0.f, 0.f, 0.f};
            for (int c = 0; c < S.nq; c++) {
                const GAS f32x2* __restrict__ in =
                    reinterpret_cast<const GAS f32x2*>(gld(S.in)) + (size_t)c * S.cstride;
                f32x4 ar[MT], ai[MT];
#pragma unroll
                for (int m = 0; m < MT; m++) ar[m] = ai[m] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kc = 0; kc < MT; kc++) {
                    u32x4 bhi, blo;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int z = kc * 16 + lq * 4 + j;
                        f32x2 v = {0.f, 0.f};
                        if (z < nz) v = in[(size_t)z * ncols + nc];
                        uint32_t hi, lo;
                        split2(v.x, v.y, hi, lo);
                        bhi[j] = hi;
                        blo[j] = lo;
                    }
                    const u32x4 brh = rows_of(bhi, 0), brl = rows_of(blo, 0);
                    const u32x4 bih = rows_of(bhi, 1), bil = rows_of(blo, 1);
#pragma unroll
                    for (int m = 0; m < MT; m++) {
                        const int r = m * 16 + ln, k0 = kc * 16 + lq * 4;
                        const u32x4 th = *reinterpret_cast<const u32x4*>(&zh[r][k0]);
                        const u32x4 tl = *reinterpret_cast<const u32x4*>(&zl[r][k0]);
                        ar[m] = mfma3(th, tl, brh, brl, ar[m]);
                        ai[m] = mfma3(th, tl, bih, bil, ai[m]);
                    }
                }
#pragma unroll
                for (int m = 0; m < MT; m++) {
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int p = m * 16 + lq * 4 + r;
                        const float wv = (S.w && p < nz) ? gld(S.w)[(p * S.kseg + sy) * S.nq + c] : 1.0f;
                        outr[m][r] = __builtin_fmaf(wv, ar[m][r], outr[m][r]);
                        outi[m][r] = __builtin_fmaf(wv, ai[m][r], outi[m][r]);
                    }
                }
            }
            if (!col_ok) continue;
            GAS f32x2* out = reinterpret_cast<GAS f32x2*>(gst(S.out));
#pragma unroll
            for (int m = 0; m < MT; m++) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int p = m * 16 + lq * 4 + r;
                    if (p >= nz) continue;
                    const size_t o = (size_t)p * ncols + n;
                    f32x2 v = {outr[m][r], outi[m][r]};
                    if (S.accumulate) {
                        const f32x2 old = out[o];
                        v.x += old.x;
                        v.y += old.y;
                    }
                    out[o] = v;
                }
            }
        } else {
            const GAS f32x2* __restrict__ in = reinterpret_cast<const GAS f32x2*>(gld(S.in));
            for (int c = 0; c < S.nq; c++) {
                f32x4 ar[MT], ai[MT];
#pragma unroll
                for (int m = 0; m < MT; m++) ar[m] = ai[m] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kc = 0; kc < MT; kc++) {
                    u32x4 bhi, blo;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int p = kc * 16 + lq * 4 + j;
                        f32x2 v = {0.f, 0.f};
                        if (p < nz) {
                            v = in[(size_t)p * ncols + nc];
                            if (S.w) {
                                const float wv = gld(S.w)[(p * S.kseg + sy) * S.nq + c];
                                v.x *= wv;
                                v.y *= wv;
                            }
                        }
                        uint32_t hi, lo;
                        split2(v.x, v.y, hi, lo);
                        bhi[j] = hi;
                        blo[j] = lo;
                    }
                    const u32x4 brh = rows_of(bhi, 0), brl = rows_of(blo, 0);
                    const u32x4 bih = rows_of(bhi, 1), bil = rows_of(blo, 1);
#pragma unroll
                    for (int m = 0; m < MT; m++) {
                        const int r = m * 16 + ln, k0 = kc * 16 + lq * 4;
                        const u32x4 th = *reinterpret_cast<const u32x4*>(&zh[r][k0]);
                        const u32x4 tl = *reinterpret_cast<const u32x4*>(&zl[r][k0]);
                        ar[m] = mfma3(th, tl, brh, brl, ar[m]);
                        ai[m] = mfma3(th, tl, bih, bil, ai[m]);
                    }
                }
                if (!col_ok) continue;
                GAS f32x2* out = reinterpret_cast<GAS f32x2*>(gst(S.out)) + (size_t)c * S.cstride;
#pragma unroll
                for (int m = 0; m < MT; m++) {
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int z = m * 16 + lq * 4 + r;
                        if (z >= nz) continue;
                        const size_t o = (size_t)z * ncols + n;
                        f32x2 v = {ar[m][r], ai[m][r]};
                        if (S.accumulate) {
                            const f32x2 old = out[o];
                            v.x += old.x;
                            v.y += old.y;
                        }
                        out[o] = v;
                    }
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------- reconstruction
// reconstruct_coils_full (src/kspace.rs:1301-1378) without GRAPPA: per coil the window, the centred
// inverse 2D DFT img_c = Fy . (W o K_c) . Fx (inverse_2d, :1428-1453), then the Roemer combine
// sum_c s_c img_c / max(sum_c s_c^2, 1e-12).
// A block owns (unit, 16-column x tile); each of its four waves a quarter of the coils.  The block
// stages Fy (all of it) and its x tile's columns of Fx in LDS, split hi/lo, once.  Stage 1,
// D1 = K . Fx for every 16-row ky tile of the x tile: K rows as the A operand (four consecutive kx
// per lane, straight from HBM/L2), Fx columns as B with the complex structure.  An MFMA result tile
// (rows 4*lq + r, column ln) is laid out exactly as a B operand (k = 4*lq + j, column ln), so
// stage 2, img = Fy . D1, takes D1 from the wave's own registers -- split to hi/lo once per ky tile
// -- with Fy rows as A.  Each coil's image tile is weighted by its sensitivity into the wave's
// running combine; the four coil groups' combines are summed in LDS in a fixed order
// (deterministic), reusing the Fy area.
// NT = the 16-row tiles of ny (1..8, ny <= 128), which size the register arrays.
template <int NT>
__global__ __launch_bounds__(256) void recon_units(const ReconGeom* __restrict__ gp, const ReconItem* __restrict__ items,
                                                   int nitems, int blocks_per_unit) {
    extern __shared__ uint32_t rlds[];
    const ReconGeom G = *gp;
    // XCD-aware order: workgroups go to the 8 XCDs round-robin, so blocks b, b+8, b+16, ... of a
    // group of 8*blocks_per_unit share an XCD; those are one unit's x tiles, whose re-reads of the
    // unit's k-space then hit that XCD's L2 instead of HBM
    const int grp8 = blockIdx.x / (8 * blocks_per_unit), in8 = blockIdx.x % (8 * blocks_per_unit);
    const int item = grp8 * 8 + in8 % 8, t = in8 / 8;
    if (item >= nitems) return;
    const ReconItem it = items[item];
    const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const int lq = lane >> 4, ln = lane & 15;
    const int ntx = G.nxP >> 4;
    const int nx = G.nx, ny = G.ny, nyP = NT * 16, nxP = G.nxP;
    const int fys = nyP + 4, fxs = nxP + 4;                 // LDS row strides (words)
    uint32_t* fy_h = rlds;                                    // [nyP][fys]
    uint32_t* fy_l = rlds + nyP * fys;
    const int fy_words = max(2 * nyP * fys, 4 * nyP * 16 * 2);
    uint32_t* fx_h = rlds + fy_words;                         // [16][fxs]: this x tile's rows of Fx^T
    uint32_t* fx_l = fx_h + 16 * fxs;
    for (int i = threadIdx.x; i < nyP * nyP; i += 256) {
        const int r = i / nyP, k = i % nyP;
        fy_h[r * fys + k] = gld(G.fy_hi)[i];
        fy_l[r * fys + k] = gld(G.fy_lo)[i];
    }
    for (int i = threadIdx.x; i < 16 * nxP; i += 256) {
        const int r = i / nxP, k = i % nxP;
        fx_h[r * fxs + k] = gld(G.fx_hi)[(t * 16 + r) * nxP + k];
        fx_l[r * fxs + k] = gld(G.fx_lo)[(t * 16 + r) * nxP + k];
    }
    __syncthreads();
    const int x = t * 16 + ln;
    f32x4 cr[NT], ci[NT];
#pragma unroll
    for (int i = 0; i < NT; i++) cr[i] = ci[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    const GAS f32x2* __restrict__ kall = reinterpret_cast<const GAS f32x2*>(gld(it.kspace));
    const bool vec = (nx & 1) == 0;   // rows of K start 16-byte aligned
    for (int c = grp; c < G.C; c += 4) {
        const GAS f32x2* __restrict__ K = kall + (size_t)c * ny * nx;
        u32x4 dh[NT], dl[NT];
#pragma unroll
        for (int kt = 0; kt < NT; kt++) {
            // keep each tile's loads in its own iteration: hoisting all of them costs occupancy
            __asm__ volatile("" ::: "memory");
            f32x4 ar = {0.f, 0.f, 0.f, 0.f}, ai = {0.f, 0.f, 0.f, 0.f};
            const int ky = kt * 16 + ln;
            for (int kb = 0; kb < ntx; kb++) {
                const int kx0 = kb * 16 + lq * 4;
                f32x2 v[4];
                if (ky < ny && vec && kx0 + 3 < nx) {
                    const f32x4 p0 = *reinterpret_cast<const GAS f32x4*>(K + ky * nx + kx0);
                    const f32x4 p1 = *reinterpret_cast<const GAS f32x4*>(K + ky * nx + kx0 + 2);
                    v[0] = f32x2{p0.x, p0.y}; v[1] = f32x2{p0.z, p0.w};
                    v[2] = f32x2{p1.x, p1.y}; v[3] = f32x2{p1.z, p1.w};
                } else {
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        v[j] = (ky < ny && kx0 + j < nx) ? K[ky * nx + kx0 + j] : f32x2{0.f, 0.f};
                }
                u32x4 ah, al;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (G.window && ky < ny && kx0 + j < nx) {
                        const float w = gld(G.window)[ky * nx + kx0 + j];
                        v[j].x *= w;
                        v[j].y *= w;
                    }
                    uint32_t h, l;
                    split2(v[j].x, v[j].y, h, l);
                    ah[j] = h;
                    al[j] = l;
                }
                const u32x4 bh = *reinterpret_cast<const u32x4*>(fx_h + ln * fxs + kx0);
                const u32x4 bl = *reinterpret_cast<const u32x4*>(fx_l + ln * fxs + kx0);
                ar = mfma3(ah, al, rows_of(bh, 0), rows_of(bl, 0), ar);
                ai = mfma3(ah, al, rows_of(bh, 1), rows_of(bl, 1), ai);
            }
            // D1 tile (ky = kt*16 + 4*lq + r, x) is this lane's B operand of stage 2
#pragma unroll
            for (int r = 0; r < 4; r++) {
                uint32_t h, l;
                split2(ar[r], ai[r], h, l);
                dh[kt][r] = h;
                dl[kt][r] = l;
            }
        }
#pragma unroll
        for (int yt = 0; yt < NT; yt++) {
            __asm__ volatile("" ::: "memory");
            f32x4 ir = {0.f, 0.f, 0.f, 0.f}, ii = {0.f, 0.f, 0.f, 0.f};
            const int ao = (yt * 16 + ln) * fys + lq * 4;
#pragma unroll
            for (int kb = 0; kb < NT; kb++) {
                const u32x4 ah = *reinterpret_cast<const u32x4*>(fy_h + ao + kb * 16);
                const u32x4 al = *reinterpret_cast<const u32x4*>(fy_l + ao + kb * 16);
                ir = mfma3(ah, al, rows_of(dh[kb], 0), rows_of(dl[kb], 0), ir);
                ii = mfma3(ah, al, rows_of(dh[kb], 1), rows_of(dl[kb], 1), ii);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int y = yt * 16 + lq * 4 + r;
                const float sv = gld(G.sens)[((size_t)c * nyP + y) * nxP + x];
                cr[yt][r] = __builtin_fmaf(sv, ir[r], cr[yt][r]);
                ci[yt][r] = __builtin_fmaf(sv, ii[r], ci[yt][r]);
            }
        }
    }
    // the coil groups' combines, summed in group order in the Fy area
    __syncthreads();
    float* red = reinterpret_cast<float*>(rlds);   // [4][nyP][16][2]
#pragma unroll
    for (int yt = 0; yt < NT; yt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int o = ((grp * nyP + yt * 16 + lq * 4 + r) * 16 + ln) * 2;
            red[o] = cr[yt][r];
            red[o + 1] = ci[yt][r];
        }
    __syncthreads();
    if (grp != 0) return;
    GAS f32x2* out = reinterpret_cast<GAS f32x2*>(gst(it.image));
#pragma unroll
    for (int yt = 0; yt < NT; yt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int y = yt * 16 + lq * 4 + r;
            float sr = cr[yt][r], si = ci[yt][r];
            for (int g = 1; g < 4; g++) {
                const int o = ((g * nyP + y) * 16 + ln) * 2;
                sr += red[o];
                si += red[o + 1];
            }
            if (x < nx && y < ny) {
                const float rs = gld(G.rssum)[y * nxP + x];
                out[y * nx + x] = f32x2{sr * rs, si * rs};
            }
        }
}

size_t recon_lds_bytes(int nxP, int nyP) {
    const size_t fy_words = std::max((size_t)2 * nyP * (nyP + 4), (size_t)4 * nyP * 16 * 2);
    return (fy_words + 2 * 16 * (size_t)(nxP + 4)) * 4;
}

hipError_t set_recon_lds_limit(size_t bytes) {
    hipError_t e;
#define RL(N) if ((e = hipFuncSetAttribute((const void*)recon_units<N>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes))) return e;
    RL(1) RL(2) RL(3) RL(4) RL(5) RL(6) RL(7) RL(8)
#undef RL
    return hipSuccess;
}

hipError_t launch_recon(const ReconGeom* g, const ReconItem* items, int nitems, int blocks_per_unit, int nty,
                        size_t lds, hipStream_t st) {
    if (nitems <= 0) return hipSuccess;
    const dim3 grid(((nitems + 7) / 8) * 8 * blocks_per_unit), block(256);
    switch (nty) {
    case 1: hipLaunchKernelGGL((recon_units<1>), grid, block, lds, st, g, items, nitems, blocks_per_unit); break;
    case 2: hipLaunchKernelGGL((recon_units<2>), grid, block, lds, st, g, items, nitems, blocks_per_unit); break;
    case 3: hipLaunchKernelGGL((recon_units<3>), grid, block, lds, st, g, items, nitems, blocks_per_unit); break;
    case 4: hipLaunchKernelGGL((recon_units<4>), grid, block, lds, st, g, items, nitems, blocks_per_unit); break;
    case 5: hipLaunchKernelGGL((recon_units<5>), grid, block, lds, st, g, items, nitems, blocks_per_unit); break;
    case 6: hipLaunchKernelGGL((recon_units<6>), grid, block, lds, st, g, items, nitems, blocks_per_unit); break;
    case 7: hipLaunchKernelGGL((recon_units<7>), grid, block, lds, st, g, items, nitems, blocks_per_unit); break;
    case 8: hipLaunchKernelGGL((recon_units<8>), grid, block, lds, st, g, items, nitems, blocks_per_unit); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ----------------------------------------------------------------------------- epilogues
// Unacquired rows are exactly zero (src/kspace.rs:839-843).  grid.y = unit.
__global__ void zero_rows(const DevUnit* __restrict__ units, const DevGroup* __restrict__ groups, int unit0) {
    const DevUnit& u = units[unit0 + blockIdx.y];
    const DevGroup& gr = groups[u.group];
    const int nx = gr.nx;
    const int total = gr.C * gr.n_zero_rows * nx;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int kx = t % nx, r = (t / nx) % gr.n_zero_rows, c = t / (nx * gr.n_zero_rows);
        u.out[((size_t)c * gr.ny + gr.zero_rows[r]) * nx + kx] = make_float2(0.f, 0.f);
    }
}

// SplitMix64 (src/kspace.rs:60-77); counter-based: draw n of state s is mix(s + n*gamma).
__device__ __forceinline__ uint64_t splitmix_at(uint64_t s, uint64_t n) {
    uint64_t z = s + n * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double unit_of(uint64_t v) { return ((double)(v >> 11) + 0.5) / 9007199254740992.0; }

// Spikes then thermal noise on acquired samples (src/kspace.rs:997-1039).  One block per
// (unit, coil): blockIdx.x = coil, blockIdx.y = unit.  Acquired samples in index order are the
// acquired rows (ascending ky) times nx.
__global__ __launch_bounds__(256) void spikes_noise(const DevUnit* __restrict__ units,
                                                    const DevGroup* __restrict__ groups, int unit0,
                                                    const int32_t* __restrict__ rows_by_group) {
    const DevUnit& u = units[unit0 + blockIdx.y];
    const DevGroup& gr = groups[u.group];
    const int c = blockIdx.x;
    if (c >= gr.C) return;
    const int nx = gr.nx, Ly = gr.Ly;
    const int32_t* rows = rows_by_group + u.group * 256;  // ascending acquired ky
    float2* K = u.out + (size_t)c * gr.ny * nx;
    const int nacq = Ly * nx;
    const uint64_t coil = u.coil0 + c;
    if (gr.n_spikes > 0) {
        __shared__ float best_m[256];
        __shared__ int best_i[256];
        float bm = -1.0f;
        int bi = 0x7fffffff;
        for (int t = threadIdx.x; t < gr.ny * nx; t += blockDim.x) {
            float2 v = K[t];
            float m = sqrtf(v.x * v.x + v.y * v.y);
            if (m > bm) { bm = m; bi = t; }
        }
        best_m[threadIdx.x] = bm;
        best_i[threadIdx.x] = bi;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = -1.0f;
            int idx = 0;
            for (int t = 0; t < (int)blockDim.x; t++)
                if (best_m[t] > m || (best_m[t] == m && best_i[t] < idx)) { m = best_m[t]; idx = best_i[t]; }
            float2 peak = m > 0.0f ? K[idx] : make_float2(0.f, 0.f);
            float2 spike = make_float2((float)((double)peak.x * gr.spike_amplitude),
                                       (float)((double)peak.y * gr.spike_amplitude));
            if (nacq > 0) {
                uint64_t s = u.seed ^ 0xA5A5123456789ABCULL;
                for (int k = 0; k < gr.n_spikes; k++) {
                    uint64_t v = splitmix_at(s, (uint64_t)k + 1);
                    int j = (int)(v % (uint64_t)nacq);
                    K[(size_t)rows[j / nx] * nx + (j % nx)] = spike;
                }
            }
        }
        __syncthreads();
    }
    if (gr.noise) {
        const uint64_t s0 = (u.seed ^ (coil * 0x9E3779B9ULL)) | 1ULL;
        for (int j = threadIdx.x; j < nacq; j += blockDim.x) {
            const uint64_t b = 4ull * (uint64_t)j;
            double u1 = unit_of(splitmix_at(s0, b + 1)), u2 = unit_of(splitmix_at(s0, b + 2));
            double u3 = unit_of(splitmix_at(s0, b + 3)), u4 = unit_of(splitmix_at(s0, b + 4));
            double gre = sqrt(-2.0 * log(u1)) * cos(kTau * u2);
            double gim = sqrt(-2.0 * log(u3)) * cos(kTau * u4);
            float2& k = K[(size_t)rows[j / nx] * nx + (j % nx)];
            k.x = (float)((double)k.x + gre * gr.noise_sigma);
            k.y = (float)((double)k.y + gim * gr.noise_sigma);
        }
    }
}

// Accuracy probe of the phase evaluation (used by the GPU tests): sin/cos of 2*pi*cyc.
__global__ void probe_sincos(const double* __restrict__ cyc, float* __restrict__ s, float* __restrict__ c, int n,
                             int mode) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mode == 0) sincos_turns(cyc[i], s[i], c[i]);
    else if (mode == 1) sincos_turns_poly(cyc[i], s[i], c[i]);
    else sincos_turns_hw(cyc[i], s[i], c[i]);
}

}  // namespace mrs

// ----------------------------------------------------------------------------- launch wrappers
namespace mrs {

// `maxtiles`: the largest number of 32 x 32 tiles of any slice's padded grid.
// list = the slices to prepare (those that own a plane set: the others read an owner's planes)
// maxq: the most compartments of any slice in the plan.
hipError_t launch_prep(DevSlice* slices, const DevGroup* groups, const int32_t* list, int nlist, int maxtiles, int mode,
                       int maxq, hipStream_t st) {
    if (nlist <= 0) return hipSuccess;
    dim3 grid(maxtiles, nlist);
    if (maxq <= 1)
        hipLaunchKernelGGL(prep_slices<1>, grid, dim3(256), 0, st, slices, groups, list, nlist, mode);
    else if (maxq == 2)
        hipLaunchKernelGGL(prep_slices<2>, grid, dim3(256), 0, st, slices, groups, list, nlist, mode);
    else if (maxq == 3)
        hipLaunchKernelGGL(prep_slices<3>, grid, dim3(256), 0, st, slices, groups, list, nlist, mode);
    else
        hipLaunchKernelGGL(prep_slices<MRS_MAX_COMP>, grid, dim3(256), 0, st, slices, groups, list, nlist, mode);
    return hipGetLastError();
}

template <int Q, int RB, int GP, bool CPLX, bool EDDY, int V = 1>
static hipError_t ylaunch(dim3 grid, hipStream_t st, const DevUnit* u, const DevSlice* s, const DevGroup* g,
                          const WorkItem* w, const int32_t* usets, int xt) {
    hipLaunchKernelGGL((ystage_mfma<Q, RB, GP, CPLX, EDDY, V>), grid, dim3(256), 0, st, u, s, g, w, usets, xt);
    return hipGetLastError();
}

// Instantiated (RB, GP) pairs: the accumulators RB*GP*4 VGPRs stay within budget.  Unit sets (V > 1)
// exist for one row block (up to 8 coils), four line tiles and no eddy term.
template <int Q, bool CPLX, bool EDDY>
static hipError_t ylaunch_rb(int rb, int gp, int nv, dim3 grid, hipStream_t st, const DevUnit* u, const DevSlice* s,
                             const DevGroup* g, const WorkItem* w, const int32_t* usets, int xt) {
    if (nv > 1) {
        if constexpr (!EDDY) {
            if (rb == 1 && gp >= 4) {
                if (nv == 2) return ylaunch<Q, 1, 4, CPLX, EDDY, 2>(grid, st, u, s, g, w, usets, xt);
                // four units' images and accumulators fit without spilling for up to three image planes
                if constexpr ((CPLX ? 2 * Q : Q) <= 3) {
                    if (nv == 4) return ylaunch<Q, 1, 4, CPLX, EDDY, 4>(grid, st, u, s, g, w, usets, xt);
                }
            }
        }
        return hipErrorInvalidValue;
    }
    switch (rb) {
    case 1:
        if (gp >= 4) return ylaunch<Q, 1, 4, CPLX, EDDY>(grid, st, u, s, g, w, usets, xt);
        if (gp == 2) return ylaunch<Q, 1, 2, CPLX, EDDY>(grid, st, u, s, g, w, usets, xt);
        return ylaunch<Q, 1, 1, CPLX, EDDY>(grid, st, u, s, g, w, usets, xt);
    case 2:
        if (gp >= 4) return ylaunch<Q, 2, 4, CPLX, EDDY>(grid, st, u, s, g, w, usets, xt);
        if (gp >= 2) return ylaunch<Q, 2, 2, CPLX, EDDY>(grid, st, u, s, g, w, usets, xt);
        return ylaunch<Q, 2, 1, CPLX, EDDY>(grid, st, u, s, g, w, usets, xt);
    case 4:
        // 16 accumulator tiles fit without spilling for up to two compartments
        if constexpr (Q <= 2) {
            if (gp >= 4) return ylaunch<Q, 4, 4, CPLX, EDDY>(grid, st, u, s, g, w, usets, xt);
        }
        return ylaunch<Q, 4, 1, CPLX, EDDY>(grid, st, u, s, g, w, usets, xt);
    case 8: return ylaunch<Q, 8, 1, CPLX, EDDY>(grid, st, u, s, g, w, usets, xt);
    default: return hipErrorInvalidValue;
    }
}

template <bool CPLX, bool EDDY>
static hipError_t ylaunch_q(int Q, int RB, int GP, int nv, dim3 grid, hipStream_t st, const DevUnit* u,
                            const DevSlice* s, const DevGroup* g, const WorkItem* w, const int32_t* usets, int xt) {
    switch (Q) {
    case 1: return ylaunch_rb<1, CPLX, EDDY>(RB, GP, nv, grid, st, u, s, g, w, usets, xt);
    case 2: return ylaunch_rb<2, CPLX, EDDY>(RB, GP, nv, grid, st, u, s, g, w, usets, xt);
    case 3: return ylaunch_rb<3, CPLX, EDDY>(RB, GP, nv, grid, st, u, s, g, w, usets, xt);
    case 4: return ylaunch_rb<4, CPLX, EDDY>(RB, GP, nv, grid, st, u, s, g, w, usets, xt);
    default: return hipErrorInvalidValue;
    }
}

// nv = units per work item (1, or 2 / 4 for sets of one slice's units listed in `usets`).
hipError_t launch_ystage_mfma(int Q, int RB, int GP, int nv, bool cplx, bool eddy, int nwork, int xt, const DevUnit* u,
                              const DevSlice* s, const DevGroup* g, const WorkItem* w, const int32_t* usets,
                              hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    dim3 grid(nwork);
    if (cplx) return eddy ? ylaunch_q<true, true>(Q, RB, GP, nv, grid, st, u, s, g, w, usets, xt)
                          : ylaunch_q<true, false>(Q, RB, GP, nv, grid, st, u, s, g, w, usets, xt);
    return eddy ? ylaunch_q<false, true>(Q, RB, GP, nv, grid, st, u, s, g, w, usets, xt)
                : ylaunch_q<false, false>(Q, RB, GP, nv, grid, st, u, s, g, w, usets, xt);
}

// nw = 4: mt = 1..4; nw = 8: mt = 3..6.
hipError_t launch_xstage_mfma(int mt, int nw, int nparts, size_t lds, int nwork, int nblk, const DevUnit* u,
                              const DevGroup* g, const WorkItem* w, hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    const dim3 grid(std::max(1, std::min(nwork, nblk)));
    if (nw == 4) {
        switch (mt) {
        case 1: hipLaunchKernelGGL((xstage_mfma<1>), grid, dim3(256), lds, st, u, g, w, nwork, nparts); break;
        case 2: hipLaunchKernelGGL((xstage_mfma<2>), grid, dim3(256), lds, st, u, g, w, nwork, nparts); break;
        case 3: hipLaunchKernelGGL((xstage_mfma<3>), grid, dim3(256), lds, st, u, g, w, nwork, nparts); break;
        case 4: hipLaunchKernelGGL((xstage_mfma<4>), grid, dim3(256), lds, st, u, g, w, nwork, nparts); break;
        default: return hipErrorInvalidValue;
        }
    } else if (nw == 8) {
        switch (mt) {
        case 3: hipLaunchKernelGGL((xstage_mfma<3, 8>), grid, dim3(512), lds, st, u, g, w, nwork, nparts); break;
        case 4: hipLaunchKernelGGL((xstage_mfma<4, 8>), grid, dim3(512), lds, st, u, g, w, nwork, nparts); break;
        case 5: hipLaunchKernelGGL((xstage_mfma<5, 8>), grid, dim3(512), lds, st, u, g, w, nwork, nparts); break;
        case 6: hipLaunchKernelGGL((xstage_mfma<6, 8>), grid, dim3(512), lds, st, u, g, w, nwork, nparts); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t set_xstage_lds_limit(size_t bytes) {
    const void* fns[] = {(const void*)xstage_mfma<1>,    (const void*)xstage_mfma<2>,    (const void*)xstage_mfma<3>,
                         (const void*)xstage_mfma<4>,    (const void*)xstage_mfma<3, 8>, (const void*)xstage_mfma<4, 8>,
                         (const void*)xstage_mfma<5, 8>, (const void*)xstage_mfma<6, 8>};
    for (const void* f : fns) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_plain_unit(int unit, int C, int LyP, int snx, int nx, const DevUnit* u, const DevSlice* s,
                             const DevGroup* g, hipStream_t st) {
    auto gy = [&](int ys) { return dim3(snx * ((LyP + 256 / ys - 1) / (256 / ys))); };
    // y-parts per block: 32 up to 8 coils (a 128^2 slice fills 1024 blocks), 16 up to 16, 4 above
    // (the LDS of the partial sums is 2 KiB per coil slot whatever the split)
    if (C <= 8) hipLaunchKernelGGL((ystage_plain<8, 32>), gy(32), dim3(256), 0, st, u, s, g, unit);
    else if (C <= 16) hipLaunchKernelGGL((ystage_plain<16, 16>), gy(16), dim3(256), 0, st, u, s, g, unit);
    else if (C <= 32) hipLaunchKernelGGL((ystage_plain<32, 4>), gy(4), dim3(256), 0, st, u, s, g, unit);
    else hipLaunchKernelGGL((ystage_plain<64, 4>), gy(4), dim3(256), 0, st, u, s, g, unit);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int nxt = LyP * C * nx;
    hipLaunchKernelGGL((xstage_plain<16>), dim3((nxt + 15) / 16), dim3(256), 0, st, u, g, unit);
    return hipGetLastError();
}

hipError_t launch_zero_rows(int unit0, int nunits, const DevUnit* u, const DevGroup* g, hipStream_t st) {
    if (nunits <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_rows, dim3(64, nunits), dim3(256), 0, st, u, g, unit0);
    return hipGetLastError();
}

// Spikes and noise of a 3D train's acquired k-space (partitions_to_slices, src/kspace3d.rs:235-264),
// one block per (partition, coil), layout [p][coil][ky][kx]: the stream seed of partition p of
// volume g is pseed = ((g * 0x100_0001 + p) * 0x9E37) ^ seed ^ SEED_SALT_3D; spikes overwrite
// acquired samples picked by Rng(pseed ^ 0xA5A5_1234_5678_9ABC ^ q * 0xC2B2_AE3D_27D4_EB4F) with
// the partition's first peak times the amplitude; then noise on every acquired sample in index
// order from Rng((pseed ^ q * 0x9E37_79B9) | 1), sample j using draws 4j+1..4j+4.
__global__ __launch_bounds__(256) void partition_noise(float2* __restrict__ kall, int C, int ny, int nx,
                                                       const int32_t* __restrict__ rows, int nrows, uint64_t g,
                                                       uint64_t seed, int n_spikes, double amp, double sigma) {
    const int p = blockIdx.y, q = blockIdx.x;
    float2* K = kall + ((size_t)p * C + q) * ny * nx;
    const uint64_t pseed = ((g * 0x1000001ULL + (uint64_t)p) * 0x9E37ULL) ^ seed ^ 0x334452454144ULL;
    const int nacq = nrows * nx;
    if (n_spikes > 0) {
        __shared__ float best_m[256];
        __shared__ int best_i[256];
        float bm = -1.0f;
        int bi = 0x7fffffff;
        for (int t = threadIdx.x; t < ny * nx; t += blockDim.x) {
            const float2 v = K[t];
            const float m = sqrtf(v.x * v.x + v.y * v.y);
            if (m > bm) { bm = m; bi = t; }
        }
        best_m[threadIdx.x] = bm;
        best_i[threadIdx.x] = bi;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = -1.0f;
            int idx = 0;
            for (int t = 0; t < (int)blockDim.x; t++)
                if (best_m[t] > m || (best_m[t] == m && best_i[t] < idx)) { m = best_m[t]; idx = best_i[t]; }
            const float2 peak = m > 0.0f ? K[idx] : make_float2(0.f, 0.f);
            const float2 spike = make_float2((float)((double)peak.x * amp), (float)((double)peak.y * amp));
            if (nacq > 0) {
                const uint64_t s = pseed ^ 0xA5A5123456789ABCULL ^ ((uint64_t)q * 0xC2B2AE3D27D4EB4FULL);
                for (int k = 0; k < n_spikes; k++) {
                    const int j = (int)(splitmix_at(s, (uint64_t)k + 1) % (uint64_t)nacq);
                    K[(size_t)rows[j / nx] * nx + (j % nx)] = spike;
                }
            }
        }
        __syncthreads();
    }
    if (sigma > 0.0) {
        const uint64_t s0 = (pseed ^ ((uint64_t)q * 0x9E3779B9ULL)) | 1ULL;
        for (int j = threadIdx.x; j < nacq; j += blockDim.x) {
            const uint64_t b = 4ull * (uint64_t)j;
            const double u1 = unit_of(splitmix_at(s0, b + 1)), u2 = unit_of(splitmix_at(s0, b + 2));
            const double u3 = unit_of(splitmix_at(s0, b + 3)), u4 = unit_of(splitmix_at(s0, b + 4));
            const double gre = sqrt(-2.0 * log(u1)) * cos(kTau * u2);
            const double gim = sqrt(-2.0 * log(u3)) * cos(kTau * u4);
            float2& k = K[(size_t)rows[j / nx] * nx + (j % nx)];
            k.x = (float)((double)k.x + gre * sigma);
            k.y = (float)((double)k.y + gim * sigma);
        }
    }
}

hipError_t launch_partition_noise(float2* k, int nz, int C, int ny, int nx, const int32_t* rows, int nrows, uint64_t g,
                                  uint64_t seed, int n_spikes, double amp, double sigma, hipStream_t st) {
    if (nz <= 0 || C <= 0) return hipSuccess;
    hipLaunchKernelGGL(partition_noise, dim3(C, nz), dim3(256), 0, st, k, C, ny, nx, rows, nrows, g, seed, n_spikes,
                       amp, sigma);
    return hipGetLastError();
}

// The tail of acquire_volumes (src/kspace.rs:2036-2060, 1840-1846): combined images [g][vox] of a
// series (vox = x + nx*(y + ny*z)) -> f32, plus the image-space noise of SD sigma[vox] from
// Rng((seed ^ g*0x9E37_79B9_7F4A_7C15 ^ vox*0xBF58_476D_1CE4_E5B9) | 1), two draws per component
// (re then im), written to the entry point's layout vox*nv + g.  mode 0: real and imaginary parts;
// mode 1: magnitude sqrt(re*re + im*im) and phase atan2(im, re) in f32, without contraction.
__global__ __launch_bounds__(256) void series_epilogue(const float2* __restrict__ img, int nv, int nvox,
                                                       const float* __restrict__ sigma, uint64_t seed, int mode,
                                                       float* __restrict__ o0, float* __restrict__ o1) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nv * nvox) return;
    const int g = i / nvox, vox = i - g * nvox;
    const float2 c = img[i];
    float re = c.x, im = c.y;
    if (sigma) {
        const double sd = (double)sigma[vox];
        if (sd > 0.0) {
            const uint64_t s0 = (seed ^ ((uint64_t)g * 0x9E3779B97F4A7C15ULL) ^ ((uint64_t)vox * 0xBF58476D1CE4E5B9ULL)) | 1ULL;
            const double u1 = unit_of(splitmix_at(s0, 1)), u2 = unit_of(splitmix_at(s0, 2));
            const double u3 = unit_of(splitmix_at(s0, 3)), u4 = unit_of(splitmix_at(s0, 4));
            re += (float)(sqrt(-2.0 * log(u1)) * cos(kTau * u2) * sd);
            im += (float)(sqrt(-2.0 * log(u3)) * cos(kTau * u4) * sd);
        }
    }
    const size_t o = (size_t)vox * nv + g;
    if (mode == 0) {
        o0[o] = re;
        o1[o] = im;
    } else {
        o0[o] = sqrtf(re * re + im * im);
        o1[o] = atan2f(im, re);
    }
}

hipError_t launch_series_epilogue(const float2* img, int nv, int nvox, const float* sigma, uint64_t seed, int mode,
                                  float* o0, float* o1, hipStream_t st) {
    const long n = (long)nv * nvox;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(series_epilogue, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, img, nv, nvox, sigma, seed,
                       mode, o0, o1);
    return hipGetLastError();
}

hipError_t launch_spikes_noise(int unit0, int nunits, int maxC, const DevUnit* u, const DevGroup* g,
                               const int32_t* rows, hipStream_t st) {
    if (nunits <= 0) return hipSuccess;
    hipLaunchKernelGGL(spikes_noise, dim3(maxC, nunits), dim3(256), 0, st, u, g, unit0, rows);
    return hipGetLastError();
}


template <int Q, bool CPLX>
static hipError_t yadj_c(int C, dim3 grid, hipStream_t st, const DevUnit* u, const DevSlice* s, const DevGroup* g,
                         const VjpUnit* v, int unit0) {
    if (C <= 8) hipLaunchKernelGGL((yadjoint<Q, CPLX, 8>), grid, dim3(256), 0, st, u, s, g, v, unit0);
    else if (C <= 16) hipLaunchKernelGGL((yadjoint<Q, CPLX, 16>), grid, dim3(256), 0, st, u, s, g, v, unit0);
    else if (C <= 32) hipLaunchKernelGGL((yadjoint<Q, CPLX, 32>), grid, dim3(256), 0, st, u, s, g, v, unit0);
    else hipLaunchKernelGGL((yadjoint<Q, CPLX, 64>), grid, dim3(256), 0, st, u, s, g, v, unit0);
    return hipGetLastError();
}

template <bool CPLX>
static hipError_t yadj_q(int Q, int C, dim3 grid, hipStream_t st, const DevUnit* u, const DevSlice* s,
                         const DevGroup* g, const VjpUnit* v, int unit0) {
    switch (Q) {
    case 1: return yadj_c<1, CPLX>(C, grid, st, u, s, g, v, unit0);
    case 2: return yadj_c<2, CPLX>(C, grid, st, u, s, g, v, unit0);
    case 3: return yadj_c<3, CPLX>(C, grid, st, u, s, g, v, unit0);
    case 4: return yadj_c<4, CPLX>(C, grid, st, u, s, g, v, unit0);
    default: return hipErrorInvalidValue;
    }
}

template <int Q, bool CPLX, bool EDDY>
static hipError_t yadjm_c(int C, int LyP, dim3 grid, hipStream_t st, const DevUnit* u, const DevSlice* s,
                          const DevGroup* g, const VjpUnit* v, int unit0) {
    if (C <= 16)
        hipLaunchKernelGGL((yadjoint_mfma<Q, CPLX, 1, EDDY>), grid, dim3(256), yadjoint_mfma_lds(LyP, 1), st, u, s, g, v,
                           unit0);
    else if (C <= 32)
        hipLaunchKernelGGL((yadjoint_mfma<Q, CPLX, 2, EDDY>), grid, dim3(256), yadjoint_mfma_lds(LyP, 2), st, u, s, g, v,
                           unit0);
    else
        hipLaunchKernelGGL((yadjoint_mfma<Q, CPLX, 4, EDDY>), grid, dim3(256), yadjoint_mfma_lds(LyP, 4), st, u, s, g, v,
                           unit0);
    return hipGetLastError();
}

template <int Q, bool CPLX, bool EDDY>
static hipError_t yadjm_lds_limit(int bytes) {
    hipError_t e;
    const auto attr = hipFuncAttributeMaxDynamicSharedMemorySize;
    if ((e = hipFuncSetAttribute((const void*)yadjoint_mfma<Q, CPLX, 1, EDDY>, attr, bytes))) return e;
    if ((e = hipFuncSetAttribute((const void*)yadjoint_mfma<Q, CPLX, 2, EDDY>, attr, bytes))) return e;
    return hipFuncSetAttribute((const void*)yadjoint_mfma<Q, CPLX, 4, EDDY>, attr, bytes);
}
template <int Q, bool CPLX>
static hipError_t yadjm_lds_limit(int bytes) {
    const hipError_t e = yadjm_lds_limit<Q, CPLX, false>(bytes);
    return e ? e : yadjm_lds_limit<Q, CPLX, true>(bytes);
}

template <bool CPLX, bool EDDY>
static hipError_t yadjm_q(int Q, int C, int LyP, dim3 grid, hipStream_t st, const DevUnit* u, const DevSlice* s,
                          const DevGroup* g, const VjpUnit* v, int unit0) {
    switch (Q) {
    case 1: return yadjm_c<1, CPLX, EDDY>(C, LyP, grid, st, u, s, g, v, unit0);
    case 2: return yadjm_c<2, CPLX, EDDY>(C, LyP, grid, st, u, s, g, v, unit0);
    case 3: return yadjm_c<3, CPLX, EDDY>(C, LyP, grid, st, u, s, g, v, unit0);
    case 4: return yadjm_c<4, CPLX, EDDY>(C, LyP, grid, st, u, s, g, v, unit0);
    default: return hipErrorInvalidValue;
    }
}

// The MFMA y-adjoint's LDS (lines + split Gb) reaches 151.5 KiB at 256 lines and 64 coils.
hipError_t set_yadjoint_lds_limit(size_t bytes) {
    hipError_t e;
    const int b = (int)bytes;
    if ((e = yadjm_lds_limit<1, false>(b)) || (e = yadjm_lds_limit<2, false>(b)) || (e = yadjm_lds_limit<3, false>(b)) ||
        (e = yadjm_lds_limit<4, false>(b)) || (e = yadjm_lds_limit<1, true>(b)) || (e = yadjm_lds_limit<2, true>(b)) ||
        (e = yadjm_lds_limit<3, true>(b)))
        return e;
    return yadjm_lds_limit<4, true>(b);
}

// Adjoint of units [unit0, unit0 + n) sharing one group shape.  mfma: the x-adjoint and the
// y-adjoint's coil contraction as split-BF16 MFMA (the y-adjoint needs the split coil planes), else
// FP32 on the VALU (the y-adjoint reads the FP32 planes).
// yadj: also launch the per-unit y-adjoint (off where the slice-basis y-adjoint takes the units).
hipError_t launch_vjp_units(int unit0, int n, int Q, int C, bool cplx, bool eddy, int snx, int sny, int nx, int LyP,
                            int ntab, bool mfma, bool yadj, const DevUnit* u, const DevSlice* s, const DevGroup* g,
                            const VjpUnit* v, hipStream_t st) {
    // MFMA x-adjoint tiling: 16*MT x rows per block with their twiddle rows within 64 KiB of LDS, the
    // unit's column tiles split over enough parts for about 1024 blocks per launch
    const int nkP = (nx + 15) & ~15;
    auto lds_of = [&](int mt) { return (size_t)ntab * 2 * 16 * mt * (nkP + 4) * 4; };
    int mt = 4;
    while (mt > 1 && lds_of(mt) > 64 * 1024) mt /= 2;
    const int nxb = (snx + 16 * mt - 1) / (16 * mt), ntiles = LyP * C / 16;
    // units ride on a grid dimension of at most 65535 blocks
    for (int off = 0; off < n; off += 65535) {
        const int m = std::min(65535, n - off);
        if (mfma) {
            const int np = std::max(1, std::min(std::max(1, ntiles / 4), 1024 / std::max(1, m * nxb) + 1));
            const int nwork = m * np * nxb;
            const dim3 grid((nwork + 7) / 8 * 8);
            if (mt == 4)
                hipLaunchKernelGGL((xadjoint_mfma<4>), grid, dim3(256), lds_of(4), st, u, g, v, unit0 + off, np, nxb, nwork);
            else if (mt == 2)
                hipLaunchKernelGGL((xadjoint_mfma<2>), grid, dim3(256), lds_of(2), st, u, g, v, unit0 + off, np, nxb, nwork);
            else
                hipLaunchKernelGGL((xadjoint_mfma<1>), grid, dim3(256), lds_of(1), st, u, g, v, unit0 + off, np, nxb, nwork);
        } else {
            hipLaunchKernelGGL(xadjoint, dim3((snx + 63) / 64, (LyP / 4) * ((C + XA_CG - 1) / XA_CG), m), dim3(256), 0,
                               st, u, g, v, unit0 + off);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (!yadj) continue;
        if (mfma) {
            dim3 grid(snx, m);
            if (eddy)
                e = cplx ? yadjm_q<true, true>(Q, C, LyP, grid, st, u, s, g, v, unit0 + off)
                         : yadjm_q<false, true>(Q, C, LyP, grid, st, u, s, g, v, unit0 + off);
            else
                e = cplx ? yadjm_q<true, false>(Q, C, LyP, grid, st, u, s, g, v, unit0 + off)
                         : yadjm_q<false, false>(Q, C, LyP, grid, st, u, s, g, v, unit0 + off);
        } else {
            dim3 grid(snx * ((sny + 255) / 256), m);
            e = cplx ? yadj_q<true>(Q, C, grid, st, u, s, g, v, unit0 + off)
                     : yadj_q<false>(Q, C, grid, st, u, s, g, v, unit0 + off);
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}


template <int MT>
static void zlaunch(bool adjoint, int nblocks, hipStream_t st, const ZSet* sets, const ZBlock* blocks) {
    if (adjoint)
        hipLaunchKernelGGL((zstage<MT, true>), dim3(nblocks), dim3(256), 0, st, sets, blocks);
    else
        hipLaunchKernelGGL((zstage<MT, false>), dim3(nblocks), dim3(256), 0, st, sets, blocks);
}

// Blocks are grouped by the table size MT (nz <= 16, 32, 64): one launch per MT present.
hipError_t launch_zstage(int mt, bool adjoint, const ZSet* sets, const ZBlock* blocks, int nblocks, hipStream_t st) {
    if (nblocks <= 0) return hipSuccess;
    switch (mt) {
    case 1: zlaunch<1>(adjoint, nblocks, st, sets, blocks); break;
    case 2: zlaunch<2>(adjoint, nblocks, st, sets, blocks); break;
    case 4: zlaunch<4>(adjoint, nblocks, st, sets, blocks); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int zstage_columns_per_block() { return kZCols; }

hipError_t launch_probe_sincos(const double* cyc, float* s, float* c, int n, int mode, hipStream_t st) {
    hipLaunchKernelGGL(probe_sincos, dim3((n + 255) / 256), dim3(256), 0, st, cyc, s, c, n, mode);
    return hipGetLastError();
}

}  // namespace mrs
