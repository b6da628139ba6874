// ybasis.hip -- the y-stage of the volumes of one slice, evaluated together (src/kspace.rs:940-975 per
// coil and volume in the reference).
//
// The volumes of a slice share every per-voxel quantity but their compartment images m_q, so the
// y-contraction of unit u, coil c, line l at column x is written as one product of a per-unit and a
// per-slice factor:
//   G_u[c, l] = sum_{q, y} A_u[(c), (q, y)] * B[(q, y), l]
//   A_u(c, q, y) = S_c(x, y) * m_{u,q}(x, y)                      (line-independent)
//   B(q, y, l)   = 2^{trf_l nr2_q + tau_l nri_q} * cis(2 pi cyc)  (unit-independent)
//   cyc          = base + rate * t_l + kyn_l * (y - sys - yoff)   (formed and reduced in FP64)
// B is generated once per (slice, column, chunk of 16 y) into LDS and read by every unit of the set;
// A is generated in registers once per (unit, column, y chunk, compartment) and used by all lines.
// Both are split into bf16 hi + lo and multiplied on v_mfma_f32_16x16x32_bf16 as hi*hi + hi*lo +
// lo*hi with FP32 accumulation.
//
// Tile layout.  A 16-row tile holds two units x 8 coils, row = 8 * (unit of the pair) + coil, so
// every lane forms its own row's natural complex A (re | im << 16 per y) with no duplicated work.
// The complex product rides on B: with K = (y, re/im),
//   B_re[(y, re), l] = B_re,  B_re[(y, im), l] = -B_im   ->  Re G
//   B_im[(y, re), l] = B_im,  B_im[(y, im), l] =  B_re   ->  Im G
// and the B_im form is the B_re form with its halves swapped and the low half's sign flipped.
//
// Work decomposition.  Block = 4 waves = one (set, column x).  A set holds up to kBasisMax units of
// one slice, in pairs (2p, 2p + 1); wave w takes the pairs w, w + 4, w + 8 (NPF of them) and all four
// 16-line tiles (64 acquired lines at most), so each A is formed once and used by every line tile;
// a last round of one or two pairs is split over the waves by line tile (template LAST).
// Per 16-y chunk:
//   1. barrier; B (generated in registers after the previous chunk's contraction) is written to LDS
//      in the MFMA B-operand layout (both forms, hi and lo), with the units' images of the chunk
//      (loaded into registers one chunk ahead);
//   2. barrier; each wave contracts its pairs against its line tiles, with the coil maps (FP32,
//      [c][x][y]) in registers loaded one chunk ahead;
//   3. each wave generates B of the next chunk for 4 of its y (one per lane's line) from the
//      column's static planes, staged in LDS once per block.
// Two blocks share a CU (at most 256 VGPRs per wave), so one block's B generation overlaps the
// other's matrix work.
#include "devcommon.h"

namespace mrs {

template <int Q, bool CPLX, int NPF>
struct BasisSmem {
    static constexpr int NU = 8 * NPF;        // units of a set (pairs wave + 4 i of 4 waves)
    static constexpr int CM = CPLX ? 2 : 1;   // floats per image value
    static constexpr int IMG_DW = NU * Q * 16 * CM;   // image dwords of one chunk
    u32x4 b[Q][4][2][2][64];                  // [q][line tile][form re/im][hi/lo][lane]
    float img[IMG_DW];                        // [unit][q][y 16][part] of the current chunk
    double base[kBasisMaxY];                  // the column's static planes, y = 0 .. snyP-1
    float rate[kBasisMaxY];
    float nr2[Q][kBasisMaxY];
    float nri[Q][kBasisMaxY];
};

template <int Q, bool CPLX, int NPF, int LAST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void ystage_basis(const DevUnit* __restrict__ units, const DevSlice* __restrict__ slices,
                  const DevGroup* __restrict__ groups, const WorkItem* __restrict__ work,
                  const int32_t* __restrict__ bsets) {
    using L = BasisSmem<Q, CPLX, NPF>;
    constexpr int NU = L::NU, CM = L::CM;
    constexpr int NSLOT = (L::IMG_DW + 255) / 256;   // image values per thread and chunk
    __shared__ __attribute__((aligned(16))) L sm;

    const WorkItem wi = work[blockIdx.x];
    const int32_t* set = bsets + (size_t)wi.unit * kBasisStride;
    const int nu = set[0];
    const int x = wi.tile;
    const DevUnit& u0 = units[set[1]];
    const DevGroup& gr = groups[u0.group];
    const DevSlice& sl = slices[u0.slice];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lq = lane >> 4, ln = lane & 15;
    const int snx = gr.snx, sny = gr.sny;
    const int snyP = (sny + 15) & ~15, snxP = (snx + 15) & ~15;
    const int C = gr.C, LyP = gr.LyP, nt = LyP >> 4;
    const int nch = snyP >> 4;
    const uint32_t P = (uint32_t)(snx * sny);
    const double sys_yoff = gr.sys_yoff;
    // Pairs per round: wave w takes pair w + 4 i in rounds i < NPF - 1.  LAST = the line tiles of
    // the last round per wave: 4 = the same as the others (pair w + 4 (NPF - 1)); when the last round
    // holds two pairs (LAST = 2: wave w takes tiles 2 (w & 1), +1 of pair 4 (NPF - 1) + (w >> 1)) or
    // one (LAST = 1: tile w of it), its (pair, line tile) products are split over the four waves, so
    // no wave multiplies a pair slot past the set's units: the 10 pairs of 20 volumes cost 10 pairs'
    // MFMA work, not 12.  Each wave forms A of its last-round pair itself.  The host picks LAST from
    // the set's size, with 64 padded lines (four tiles).
    static_assert(LAST == 4 || LAST == 2 || LAST == 1, "tiles of the last round per wave");
    constexpr int NMAIN = LAST == 4 ? NPF : NPF - 1;   // rounds over all four tiles
    const int last_pair = 4 * (NPF - 1) + (LAST == 4 ? wave : (LAST == 2 ? (wave >> 1) : 0));
    const int t_lo = LAST == 4 ? 0 : (LAST == 2 ? 2 * (wave & 1) : wave);
    auto pair_of = [&](int i) { return i < NPF - 1 ? wave + 4 * i : last_pair; };

    // this lane's line of the B generation (lane = line; LyP <= 64)
    double t_s = 0.0, kyn = 0.0;
    float trf = 0.0f, tau = 0.0f;
    bool ok = false;
    if (lane < LyP) {
        const LineRec& Lr = gr.lines[lane];
        t_s = Lr.t_s;
        kyn = Lr.kyn;
        trf = Lr.trf;
        tau = Lr.tau;
        ok = Lr.ky >= 0;
    }

    // Image values: thread slot k holds dword d = 256 k + tid of a chunk in the LDS order
    // (unit, q, y, part).  Loaded into registers one chunk ahead, written to LDS after the chunk's
    // first barrier.
    const GAS float* isrc[NSLOT];
    int iy[NSLOT];
#pragma unroll
    for (int k = 0; k < NSLOT; k++) {
        const int d = 256 * k + tid;
        isrc[k] = nullptr;
        iy[k] = 0;
        const int uu = d / (Q * 16 * CM);
        if (d < L::IMG_DW && uu < nu) {   // slots past the set's units stay unset (rows never stored)
            const int r = d % (Q * 16 * CM);
            const int q = r / (16 * CM), yy = (r / CM) & 15, pt = r % CM;
            const float* img = (const float*)units[set[1 + uu]].images;
            isrc[k] = gld(img) + ((size_t)q * P + x) * CM + pt;
            iy[k] = yy;
        }
    }
    auto load_images = [&](int ch, float (&v)[NSLOT]) {
#pragma unroll
        for (int k = 0; k < NSLOT; k++) {
            v[k] = 0.0f;
            if (isrc[k]) v[k] = isrc[k][(size_t)snx * CM * min(ch * 16 + iy[k], sny - 1)];
        }
    };

    // coil maps of this lane's row (coil ln & 7) at the chunk's 4 y of the lane: FP32, [c][x][y]
    const int coil = ln & 7;
    const bool has_coil = coil < C;
    const GAS float2* sbase = gld(sl.s32) + ((size_t)(has_coil ? coil : 0) * snxP + x) * snyP + 4 * lq;
    auto load_s = [&](int ch, f32x4& s01, f32x4& s23) {
        s01 = s23 = f32x4{0.f, 0.f, 0.f, 0.f};
        if (has_coil) {
            const GAS f32x4* p = (const GAS f32x4*)(sbase + ch * 16);
            s01 = p[0];
            s23 = p[1];
        }
    };

    // B of this wave's quad of y (ch * 16 + 4 wave .. +3) at the lane's line, split, in the B_re form,
    // from the column's static planes in LDS (wave-uniform addresses: broadcast reads)
    auto gen_b = [&](int ch, u32x4 (&hre)[Q], u32x4 (&lre)[Q]) {
        const int yq = ch * 16 + 4 * wave;
        const f64x2 b01 = *reinterpret_cast<const f64x2*>(&sm.base[yq]);
        const f64x2 b23 = *reinterpret_cast<const f64x2*>(&sm.base[yq + 2]);
        const f32x4 r4 = *reinterpret_cast<const f32x4*>(&sm.rate[yq]);
        const double bb[4] = {b01.x, b01.y, b23.x, b23.y};
        f32x4 a2[Q], ai[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            a2[q] = *reinterpret_cast<const f32x4*>(&sm.nr2[q][yq]);
            ai[q] = *reinterpret_cast<const f32x4*>(&sm.nri[q][yq]);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const double yy = (double)(yq + j) - sys_yoff;
            double cyc = __builtin_fma(kyn, yy, bb[j]);
            cyc = __builtin_fma((double)r4[j], t_s, cyc);
            float sn, cs;
            sincos_turns(cyc, sn, cs);
#pragma unroll
            for (int q = 0; q < Q; q++) {
                float dq = __builtin_amdgcn_exp2f(__builtin_fmaf(trf, a2[q][j], tau * ai[q][j]));
                dq = ok ? dq : 0.0f;
                uint32_t hh, ll;
                split2(dq * cs, -(dq * sn), hh, ll);
                hre[q][j] = hh;
                lre[q][j] = ll;
            }
        }
    };

    f32x4 acc[NPF][4][2];
#pragma unroll
    for (int i = 0; i < NPF; i++)
#pragma unroll
        for (int t = 0; t < 4; t++) acc[i][t][0] = acc[i][t][1] = f32x4{0.f, 0.f, 0.f, 0.f};

    // prologue: the column's static planes into LDS (once per block), chunk 0's operands
    {
        const uint32_t o = (uint32_t)(x * snyP), PPl = (uint32_t)(snxP * snyP);
        for (int i = tid; i < snyP; i += 256) {
            sm.base[i] = gld(sl.base)[o + i];
            sm.rate[i] = gld(sl.rate)[o + i];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                sm.nr2[q][i] = gld(sl.nr2)[q * PPl + o + i];
                sm.nri[q][i] = gld(sl.nri)[q * PPl + o + i];
            }
        }
    }
    f32x4 sn01, sn23;
    float inext[NSLOT];
    load_s(0, sn01, sn23);
    load_images(0, inext);
    __syncthreads();
    u32x4 hre[Q], lre[Q];
    gen_b(0, hre, lre);

    for (int ch = 0; ch < nch; ch++) {
        // this chunk's operands (loaded one chunk ahead); the next chunk's load during this one
        const f32x4 s01 = sn01, s23 = sn23;
        float icur[NSLOT];
#pragma unroll
        for (int k = 0; k < NSLOT; k++) icur[k] = inext[k];
        if (ch + 1 < nch) {
            load_s(ch + 1, sn01, sn23);
            load_images(ch + 1, inext);
        }
        __syncthreads();   // every wave is done reading the previous chunk's B and images
        // B of the chunk (both forms) and the units' images into LDS.  Lanes past the acquired
        // lines write zero B into tiles no wave reads.
        {
            const int lt = lane >> 4, slot = wave * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < Q; q++) {
                u32x4 him, lim;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    him[j] = __builtin_amdgcn_alignbit(hre[q][j], hre[q][j], 16) ^ 0x8000u;
                    lim[j] = __builtin_amdgcn_alignbit(lre[q][j], lre[q][j], 16) ^ 0x8000u;
                }
                sm.b[q][lt][0][0][slot] = hre[q];
                sm.b[q][lt][0][1][slot] = lre[q];
                sm.b[q][lt][1][0][slot] = him;
                sm.b[q][lt][1][1][slot] = lim;
            }
        }
#pragma unroll
        for (int k = 0; k < NSLOT; k++)
            if (isrc[k]) sm.img[256 * k + tid] = icur[k];
        __syncthreads();

        // contraction: per compartment, A of each of the wave's pairs (pair wave + 4 i, units 2 pair
        // and 2 pair + 1) once, then every line tile against it (tiles past the acquired lines hold
        // zero B; their rows are not stored)
        const float sr[4] = {s01.x, s01.z, s23.x, s23.z};
        const float si[4] = {s01.y, s01.w, s23.y, s23.w};
#pragma unroll
        for (int q = 0; q < Q; q++) {
            u32x4 ah[NPF], al[NPF];
#pragma unroll
            for (int i = 0; i < NPF; i++) {
                const int ul = 2 * pair_of(i) + (ln >> 3);
                const float* mp = sm.img + ((min(ul, NU - 1) * Q + q) * 16 + 4 * lq) * CM;
                if constexpr (CPLX) {
                    const f32x4 m01 = *reinterpret_cast<const f32x4*>(mp);
                    const f32x4 m23 = *reinterpret_cast<const f32x4*>(mp + 4);
                    const float mr[4] = {m01.x, m01.z, m23.x, m23.z};
                    const float mi[4] = {m01.y, m01.w, m23.y, m23.w};
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const float ar = __builtin_fmaf(sr[j], mr[j], -(si[j] * mi[j]));
                        const float aim = __builtin_fmaf(sr[j], mi[j], si[j] * mr[j]);
                        uint32_t hh, ll;
                        split2(ar, aim, hh, ll);
                        ah[i][j] = hh;
                        al[i][j] = ll;
                    }
                } else {
                    const f32x4 m = *reinterpret_cast<const f32x4*>(mp);
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        uint32_t hh, ll;
                        split2(sr[j] * m[j], si[j] * m[j], hh, ll);
                        ah[i][j] = hh;
                        al[i][j] = ll;
                    }
                }
            }
            // line tile t's B is read while tile t - 1 is multiplied; the three split products run in
            // phases over all 2 NMAIN accumulators of the tile, so dependent MFMAs are 2 NMAIN apart
            if constexpr (NMAIN > 0) {
                u32x4 bh[2], bl[2], nbh[2], nbl[2];
#pragma unroll
                for (int f = 0; f < 2; f++) {
                    nbh[f] = sm.b[q][0][f][0][lane];
                    nbl[f] = sm.b[q][0][f][1][lane];
                }
#pragma unroll
                for (int t = 0; t < 4; t++) {
#pragma unroll
                    for (int f = 0; f < 2; f++) {
                        bh[f] = nbh[f];
                        bl[f] = nbl[f];
                    }
                    if (t + 1 < 4) {
#pragma unroll
                        for (int f = 0; f < 2; f++) {
                            nbh[f] = sm.b[q][t + 1][f][0][lane];
                            nbl[f] = sm.b[q][t + 1][f][1][lane];
                        }
                    }
#pragma unroll
                    for (int i = 0; i < NMAIN; i++)
#pragma unroll
                        for (int f = 0; f < 2; f++) acc[i][t][f] = mfma_bf16(al[i], bh[f], acc[i][t][f]);
#pragma unroll
                    for (int i = 0; i < NMAIN; i++)
#pragma unroll
                        for (int f = 0; f < 2; f++) acc[i][t][f] = mfma_bf16(ah[i], bl[f], acc[i][t][f]);
#pragma unroll
                    for (int i = 0; i < NMAIN; i++)
#pragma unroll
                        for (int f = 0; f < 2; f++) acc[i][t][f] = mfma_bf16(ah[i], bh[f], acc[i][t][f]);
                }
            }
            // the split last round: this wave's LAST tiles t_lo + k of its last-round pair
            if constexpr (LAST < 4) {
                constexpr int i = NPF - 1;
                u32x4 bh[LAST][2], bl[LAST][2];
#pragma unroll
                for (int k = 0; k < LAST; k++)
#pragma unroll
                    for (int f = 0; f < 2; f++) {
                        bh[k][f] = sm.b[q][t_lo + k][f][0][lane];
                        bl[k][f] = sm.b[q][t_lo + k][f][1][lane];
                    }
#pragma unroll
                for (int k = 0; k < LAST; k++)
#pragma unroll
                    for (int f = 0; f < 2; f++) acc[i][k][f] = mfma_bf16(al[i], bh[k][f], acc[i][k][f]);
#pragma unroll
                for (int k = 0; k < LAST; k++)
#pragma unroll
                    for (int f = 0; f < 2; f++) acc[i][k][f] = mfma_bf16(ah[i], bl[k][f], acc[i][k][f]);
#pragma unroll
                for (int k = 0; k < LAST; k++)
#pragma unroll
                    for (int f = 0; f < 2; f++) acc[i][k][f] = mfma_bf16(ah[i], bh[k][f], acc[i][k][f]);
            }
        }
        // B of the next chunk, while other waves still contract this one
        if (ch + 1 < nch) gen_b(ch + 1, hre, lre);
    }

    // G_u[x][line][coil]: lane holds rows 4 lq .. 4 lq + 3 of its tiles: unit 2 * pair + (lq >> 1) of
    // the set, coils 4 (lq & 1) .. +3, line 16 * tile + ln
#pragma unroll
    for (int i = 0; i < NPF; i++) {
        const int ul = 2 * pair_of(i) + (lq >> 1);
        if (ul >= nu) continue;
        float2* gb = units[set[1 + ul]].gbuf;
        const int c0 = 4 * (lq & 1);
        const bool split = i == NPF - 1 && LAST < 4;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (split && k >= LAST) continue;
            const int t = split ? t_lo + k : k;     // acc[i][k] holds line tile t
            if (t >= nt) continue;
            const f32x4 re = acc[i][k][0], im = acc[i][k][1];
            // stored split, the x-stage's B operand: {hi, lo} per complex value at [x][line*C + coil]
            u32x2 s[4];
#pragma unroll
            for (int r = 0; r < 4; r++) s[r] = split_g(re[r], im[r]);
            const int n0 = (t * 16 + ln) * C + c0;
            GAS u32x2* gg = reinterpret_cast<GAS u32x2*>(gst(gb));
            if (C == 8) {
                // 32-byte aligned: x*LyP*8 + n0 is a multiple of 4 complex values
                GAS u32x4* d4 = reinterpret_cast<GAS u32x4*>(gg + (size_t)x * (LyP * C) + n0);
                d4[0] = u32x4{s[0].x, s[0].y, s[1].x, s[1].y};
                d4[1] = u32x4{s[2].x, s[2].y, s[3].x, s[3].y};
            } else {
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (c0 + r < C) gg[(size_t)x * (LyP * C) + n0 + r] = s[r];
            }
        }
    }
}

// ----------------------------------------------------------------------------- launch
template <int Q, bool CPLX, int NPF>
static hipError_t blaunch_last(int last, dim3 grid, const DevUnit* u, const DevSlice* s, const DevGroup* g,
                               const WorkItem* w, const int32_t* sets, hipStream_t st) {
    switch (last) {
    case 4: hipLaunchKernelGGL((ystage_basis<Q, CPLX, NPF, 4>), grid, dim3(256), 0, st, u, s, g, w, sets); break;
    case 2: hipLaunchKernelGGL((ystage_basis<Q, CPLX, NPF, 2>), grid, dim3(256), 0, st, u, s, g, w, sets); break;
    case 1: hipLaunchKernelGGL((ystage_basis<Q, CPLX, NPF, 1>), grid, dim3(256), 0, st, u, s, g, w, sets); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int Q, bool CPLX>
static hipError_t blaunch_npw(int npf, int last, int nwork, const DevUnit* u, const DevSlice* s, const DevGroup* g,
                              const WorkItem* w, const int32_t* sets, hipStream_t st) {
    const dim3 grid(nwork);
    switch (npf) {
    case 1: return blaunch_last<Q, CPLX, 1>(last, grid, u, s, g, w, sets, st);
    case 2: return blaunch_last<Q, CPLX, 2>(last, grid, u, s, g, w, sets, st);
    case 3: return blaunch_last<Q, CPLX, 3>(last, grid, u, s, g, w, sets, st);
    default: return hipErrorInvalidValue;
    }
}

// npf = unit pairs per wave (a set holds up to 8 * npf units); last = line tiles per wave of the
// last round (4, or 2 / 1 when it holds two / one pairs of a 64-line group); Q = 1..4 compartments.
hipError_t launch_ystage_basis(int Q, bool cplx, int npw, int last, int nwork, const DevUnit* u, const DevSlice* s,
                               const DevGroup* g, const WorkItem* w, const int32_t* sets, hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    switch (Q) {
    case 1: return cplx ? blaunch_npw<1, true>(npw, last, nwork, u, s, g, w, sets, st)
                        : blaunch_npw<1, false>(npw, last, nwork, u, s, g, w, sets, st);
    case 2: return cplx ? blaunch_npw<2, true>(npw, last, nwork, u, s, g, w, sets, st)
                        : blaunch_npw<2, false>(npw, last, nwork, u, s, g, w, sets, st);
    case 3: return cplx ? blaunch_npw<3, true>(npw, last, nwork, u, s, g, w, sets, st)
                        : blaunch_npw<3, false>(npw, last, nwork, u, s, g, w, sets, st);
    case 4: return cplx ? blaunch_npw<4, true>(npw, last, nwork, u, s, g, w, sets, st)
                        : blaunch_npw<4, false>(npw, last, nwork, u, s, g, w, sets, st);
    default: return hipErrorInvalidValue;
    }
}


// ============================================================================= adjoint
// The y-adjoint of the volumes of a slice, evaluated together (the per-unit kernels are yadjoint and
// yadjoint_mfma in kernels.hip; the reference has no adjoint: this is the transpose of the y-sum of
// build_coil_kspace_timed, src/kspace.rs:940-975).
//
// With the x-adjoint's result Gb_{u,c}(x, l) of unit u, coil c, line l, and the unit-independent
//   W_q(y, l) = D_q(y, l) conj(P(y, l)),  D_q = 2^{trf_l nr2_q + tau_l nri_q},  P = cis(2 pi cyc),
// every sum the adjoint needs is a coil contraction of one product per line weight w^k:
//   R^k_{u,c}(q, y) = sum_l Gb_{u,c}(l) w^k_l W_q(y, l)      (w^0 = 1, w^1 = tau_l, w^2 = tb0_l)
//   S^k_{u,q}(y)    = sum_c conj(S_c(y)) R^k_{u,c}(q, y)
// S^0 is the image gradient, Re(conj(m_q) S^1)/T2'^2 the T2' gradient and 2 pi Im(sum_q conj(m_q) S^2)
// the fieldmap gradient.  R^k is a product of Gb [(unit, coil), line] and W^k [line, y] on
// v_mfma_f32_16x16x32_bf16 (split BF16, three products, FP32 accumulation); W^k is generated once
// per (slice, column, chunk of 16 y) for all the set's units, so the phase and the decays of a voxel
// and line are evaluated once per slice instead of once per unit.  The coil contraction runs on the
// VALU with the split coil planes (S = hi + lo).
//
// Tile layout.  A 16-row tile holds two units x 8 coils (row = 8 * unit of the pair + coil); K =
// (line, re/im) with 16 lines per step; N = 16 y of the chunk.  The complex product rides on B:
//   B_re[(l, re), y] = Re W^k,  B_re[(l, im), y] = -Im W^k   ->  Re R
//   B_im[(l, re), y] = Im W^k,  B_im[(l, im), y] =  Re W^k   ->  Im R
// Work.  Block = 4 waves = one (set, column x).  Wave w holds A (the split Gb) of the pairs
// w, w + 4, w + 8 (NPF of them) for all lines in registers; per chunk every thread generates W of one
// y and four lines, and for each needed kind writes W^k to LDS (both forms, hi and lo), after which
// the waves contract their pairs against it and fold the result into per-lane sums of the lane's
// unit, y and four coils.  A lane's accumulator rows are four coils of one unit; the lane 16 apart
// holds the other four, and one exchange completes the coil sum.

template <int Q>
struct AdjBasisSmem {
    u32x4 w[Q][4][2][2][64];   // [q][line step][form re/im][hi/lo][lane]: W^k of the chunk, B layout
    double base[kBasisMaxY];   // the column's static planes, y = 0 .. snyP-1
    float rate[kBasisMaxY];
    float nr2[Q][kBasisMaxY];
    float nri[Q][kBasisMaxY];
    LineRec lines[64];
    int32_t perm[64];          // the adjoint's K order: K position -> line (-1: none)
};

template <int Q, bool CPLX, int NPF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void yadjoint_basis(const DevUnit* __restrict__ units, const DevSlice* __restrict__ slices,
                    const DevGroup* __restrict__ groups, const VjpUnit* __restrict__ vj,
                    const WorkItem* __restrict__ work, const int32_t* __restrict__ sets) {
    constexpr int CM = CPLX ? 2 : 1;   // parts kept per sum: real images need one of each
    __shared__ __attribute__((aligned(16))) AdjBasisSmem<Q> sm;

    const WorkItem wi = work[blockIdx.x];
    const int32_t* set = sets + (size_t)wi.unit * kBasisStride;
    const int nu = set[0];
    const int x = wi.tile;
    const DevUnit& u0 = units[set[1]];
    const DevGroup& gr = groups[u0.group];
    const DevSlice& sl = slices[u0.slice];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lq = lane >> 4, ln = lane & 15;
    const int snx = gr.snx, sny = gr.sny;
    const int snyP = (sny + 15) & ~15, snxP = (snx + 15) & ~15;
    const int C = gr.C, LyP = gr.LyP;
    const int nch = snyP >> 4;
    const uint32_t P = (uint32_t)(snx * sny);
    const int np = (nu + 1) >> 1;
    // K steps of 16 lines in the adjoint's line order (the group's table: the identity but in mode 2,
    // which orders them by the sign of tb0 and splits the tb0-weighted product at ksplit)
    const int mode = gr.adj_mode, nks = gr.adj_nks, ksplit = gr.adj_ksplit;
    const float alpha = gr.adj_alpha;
    // this wave's last-round pair exists (the earlier rounds always do: the host gives a set of n
    // units the variant with NPF = ceil(ceil(n / 2) / 4))
    const bool last_ok = wave + 4 * (NPF - 1) < np;
    // the two units of each of this wave's pairs (wave-uniform: their descriptors come through scalar
    // loads, and each lane selects its unit's, lq >> 1); pairs past the set alias the last unit
    int uidx_a[NPF], uidx_b[NPF];
#pragma unroll
    for (int i = 0; i < NPF; i++) {
        const int p = wave + 4 * i;
        uidx_a[i] = set[1 + min(2 * p, nu - 1)];
        uidx_b[i] = set[1 + min(2 * p + 1, nu - 1)];
    }

    // the sums the set's units ask for (the same in every thread: block-uniform branches)
    bool need1 = false, need2 = false;
    for (int k = 0; k < nu; k++) {
        const VjpUnit& v = vj[set[1 + k]];
        need1 = need1 || v.gti != nullptr;
        need2 = need2 || v.gfmap != nullptr;
    }
    need1 = need1 && gr.do_relax;
    need2 = need2 && gr.do_dist;
    // the products to form: the image sum, then (mode 0) the tau- and the tb0-weighted ones, or
    // (modes 1, 2) one tb0-weighted product that gives both the T2' and the fieldmap sums
    const bool want1 = mode == 0 && need1, want2 = mode == 0 ? need2 : (need1 || need2);

    // the column's static planes and the line table into LDS, once per block
    {
        const uint32_t o = (uint32_t)(x * snyP), PPl = (uint32_t)(snxP * snyP);
        for (int i = tid; i < snyP; i += 256) {
            sm.base[i] = gld(sl.base)[o + i];
            sm.rate[i] = gld(sl.rate)[o + i];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                sm.nr2[q][i] = gld(sl.nr2)[q * PPl + o + i];
                sm.nri[q][i] = gld(sl.nri)[q * PPl + o + i];
            }
        }
        for (int i = tid; i < LyP; i += 256) sm.lines[i] = gr.lines[i];
        for (int i = tid; i < 64; i += 256) sm.perm[i] = gr.adj_perm[i];
    }
    __syncthreads();

    // A: the split Gb of this wave's pairs, lines 16 ks + 4 lq .. +3 of row ln (unit ln >> 3 of the
    // pair, coil ln & 7); zero for rows past the set's units or C and for steps past LyP
    u32x4 ah[NPF][4], al[NPF][4];
#pragma unroll
    for (int i = 0; i < NPF; i++) {
        const int p = wave + 4 * i;
        const int ul = 2 * p + (ln >> 3), c = ln & 7;
        const bool ok = p < np && ul < nu && c < C;
        const GAS float2* gb = gld(vj[set[1 + (ok ? ul : 0)]].gbar) + ((size_t)x * LyP) * C + (ok ? c : 0);
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int l = sm.perm[ks * 16 + 4 * lq + j];
                float2 g = make_float2(0.f, 0.f);
                if (ok && ks < nks && l >= 0) {
                    const GAS float2& v = gb[(size_t)l * C];
                    g.x = v.x;
                    g.y = v.y;
                }
                uint32_t hh, ll;
                split2(g.x, g.y, hh, ll);
                ah[i][ks][j] = hh;
                al[i][ks][j] = ll;
            }
    }

    // W generation: this thread's y = 16 ch + yl and lines 4 g4 .. 4 g4 + 3 (one B-operand slot)
    const int yl = tid & 15, g4 = tid >> 4;
    const int wks = g4 >> 2, wslot = (g4 & 3) * 16 + yl;
    const bool wact = 4 * g4 < 16 * nks;
    // the thread's four lines in the adjoint's K order (-1: none), read once
    int lkk[4];
#pragma unroll
    for (int j = 0; j < 4; j++) lkk[j] = wact ? sm.perm[4 * g4 + j] : -1;
    const double sys_yoff = gr.sys_yoff;
    // coil planes of this lane's four coils 4 (lq & 1) .. +3 (split, [x][c][y], c padded to 8)
    const int c0 = 4 * (lq & 1);

    for (int ch = 0; ch < nch; ch++) {
        // D cos, D sin of the thread's (y, lines): W = D (cos - i sin)
        float pr[Q][4], ps[Q][4];
        {
            const int y = ch * 16 + yl;
            const double bb = sm.base[y];
            const double rr = (double)sm.rate[y];
            const double yy = (double)y - sys_yoff;
            float a2[Q], ai2[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                a2[q] = sm.nr2[q][y];
                ai2[q] = sm.nri[q][y];
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int lk = lkk[j];
                const LineRec& L = sm.lines[lk >= 0 ? lk : 0];
                double cyc = __builtin_fma(L.kyn, yy, bb);
                cyc = __builtin_fma(rr, L.t_s, cyc);
                float sn, cs;
                sincos_turns(cyc, sn, cs);
                const bool ok = lk >= 0 && L.ky >= 0;
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    float d = __builtin_amdgcn_exp2f(__builtin_fmaf(L.trf, a2[q], L.tau * ai2[q]));
                    d = ok ? d : 0.0f;
                    pr[q][j] = d * cs;
                    ps[q][j] = d * sn;
                }
            }
        }
        // conj(S_c) of the lane's coils at y = 16 ch + ln (zero rows past sny and coils past C)
        float sr[4], si[4];
        {
            const int y = ch * 16 + ln;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const size_t so = ((size_t)x * 8 + c0 + r) * snyP + y;
                const uint32_t h = gld(sl.s_hi)[so], lo = gld(sl.s_lo)[so];
                sr[r] = bf16_float(h & 0xffffu) + bf16_float(lo & 0xffffu);
                si[r] = __builtin_bit_cast(float, h & 0xffff0000u) + __builtin_bit_cast(float, lo & 0xffff0000u);
            }
        }
        // the output role of this lane: unit 2 p + (lq >> 1) of pair p at y = 16 ch + ln, written by the
        // lanes with (lq & 1) == 0 once the coil sum is complete; its images and T2' (ms) at the voxel
        const int y = ch * 16 + ln;
        const bool wr_lane = (lq & 1) == 0 && y < sny;
        const uint32_t iv = (uint32_t)x + (uint32_t)snx * (uint32_t)min(y, sny - 1);
        float m_r[NPF][Q], m_i[NPF][Q], fm[NPF][CM];
#pragma unroll
        for (int i = 0; i < NPF; i++) {
            const int p = wave + 4 * i, ul = 2 * p + (lq >> 1);
            const bool ok = p < np && ul < nu;
            const GAS float* img = gld((lq >> 1) ? (const float*)units[uidx_b[i]].images : (const float*)units[uidx_a[i]].images);
#pragma unroll
            for (int q = 0; q < Q; q++) {
                if constexpr (CPLX) {
                    m_r[i][q] = ok ? img[2 * (q * P + iv)] : 0.0f;
                    m_i[i][q] = ok ? img[2 * (q * P + iv) + 1] : 0.0f;
                } else {
                    m_r[i][q] = ok ? img[q * P + iv] : 0.0f;
                    m_i[i][q] = 0.0f;
                }
            }
#pragma unroll
            for (int m = 0; m < CM; m++) fm[i][m] = 0.0f;
        }
        float ti_ms[Q];   // T2' (ms) of each compartment at the voxel, for the T2' gradient
#pragma unroll
        for (int q = 0; q < Q; q++)
            ti_ms[q] = sl.t_inhom ? gld(sl.t_inhom)[q * P + iv] : (sl.has_t_inhom ? sl.ti_uniform[q] : (float)gr.t_inhom);

        // Products over the K steps [k0, k1) of compartment q for this wave's pairs, then the coil sums
        // of the lane's four coils, completed by the lane 16 apart: ore = sum_c Re(conj(S_c) R_c),
        // oim = sum_c Im(conj(S_c) R_c), conj(S) R = (sr re + si im) + i (sr im - si re).
        auto contract = [&](int q, int k0, int k1, bool want_re, bool want_im, float (&ore)[NPF], float (&oim)[NPF]) {
            f32x4 acc[NPF][2];
#pragma unroll
            for (int i = 0; i < NPF; i++) acc[i][0] = acc[i][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                if (ks < k0 || ks >= k1) continue;   // uniform; the unrolled steps keep A in registers
                u32x4 bh[2], bl[2];
#pragma unroll
                for (int f = 0; f < 2; f++) {
                    bh[f] = sm.w[q][ks][f][0][lane];
                    bl[f] = sm.w[q][ks][f][1][lane];
                }
                if (last_ok) {
#pragma unroll
                    for (int i = 0; i < NPF; i++)
#pragma unroll
                        for (int f = 0; f < 2; f++) acc[i][f] = mfma_bf16(al[i][ks], bh[f], acc[i][f]);
#pragma unroll
                    for (int i = 0; i < NPF; i++)
#pragma unroll
                        for (int f = 0; f < 2; f++) acc[i][f] = mfma_bf16(ah[i][ks], bl[f], acc[i][f]);
#pragma unroll
                    for (int i = 0; i < NPF; i++)
#pragma unroll
                        for (int f = 0; f < 2; f++) acc[i][f] = mfma_bf16(ah[i][ks], bh[f], acc[i][f]);
                } else {
#pragma unroll
                    for (int i = 0; i < NPF - 1; i++)
#pragma unroll
                        for (int f = 0; f < 2; f++) acc[i][f] = mfma_bf16(al[i][ks], bh[f], acc[i][f]);
#pragma unroll
                    for (int i = 0; i < NPF - 1; i++)
#pragma unroll
                        for (int f = 0; f < 2; f++) acc[i][f] = mfma_bf16(ah[i][ks], bl[f], acc[i][f]);
#pragma unroll
                    for (int i = 0; i < NPF - 1; i++)
#pragma unroll
                        for (int f = 0; f < 2; f++) acc[i][f] = mfma_bf16(ah[i][ks], bh[f], acc[i][f]);
                }
            }
#pragma unroll
            for (int i = 0; i < NPF; i++) {
                float re_s = 0.f, im_s = 0.f;
                if (wave + 4 * i < np) {   // wave-uniform
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const float rr = acc[i][0][r], ri = acc[i][1][r];
                        if (want_re) re_s = __builtin_fmaf(sr[r], rr, __builtin_fmaf(si[r], ri, re_s));
                        if (want_im) im_s = __builtin_fmaf(sr[r], ri, __builtin_fmaf(-si[r], rr, im_s));
                    }
                    if (want_re) re_s += __shfl_xor(re_s, 16);
                    if (want_im) im_s += __shfl_xor(im_s, 16);
                }
                ore[i] = re_s;
                oim[i] = im_s;
            }
        };

#pragma unroll
        for (int k = 0; k < 3; k++) {
            if ((k == 1 && !want1) || (k == 2 && !want2)) continue;
            __syncthreads();   // every wave is done reading the previous W^k
            if (wact) {
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    u32x4 hre, lre, him, lim;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int lk = lkk[j];
                        const LineRec& L = sm.lines[lk >= 0 ? lk : 0];   // W is zero where lk < 0
                        const float s = k == 0 ? 1.0f : (k == 1 ? L.tau : L.tb0);
                        uint32_t hh, ll;
                        split2(pr[q][j] * s, ps[q][j] * s, hh, ll);   // B_re: (Re W, -Im W) = (D cos, D sin)
                        hre[j] = hh;
                        lre[j] = ll;
                        him[j] = __builtin_amdgcn_alignbit(hh, hh, 16) ^ 0x8000u;   // B_im: (Im W, Re W)
                        lim[j] = __builtin_amdgcn_alignbit(ll, ll, 16) ^ 0x8000u;
                    }
                    sm.w[q][wks][0][0][wslot] = hre;
                    sm.w[q][wks][0][1][wslot] = lre;
                    sm.w[q][wks][1][0][wslot] = him;
                    sm.w[q][wks][1][1][wslot] = lim;
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < Q; q++) {
                float re_a[NPF], im_a[NPF], re_b[NPF], im_b[NPF];
                // the tb0-weighted product in two parts only when the T2' sum comes from it
                const bool b_from_f = k == 2 && mode != 0 && need1;
                if (k == 0 || k == 1)
                    contract(q, 0, nks, true, CPLX, re_a, im_a);
                else if (!b_from_f)
                    contract(q, 0, nks, CPLX, true, re_a, im_a);
                else
                    contract(q, 0, ksplit, true, true, re_a, im_a);
                if (b_from_f && mode == 2)
                    contract(q, ksplit, nks, true, true, re_b, im_b);
                else
#pragma unroll
                    for (int i = 0; i < NPF; i++) re_b[i] = im_b[i] = 0.0f;
#pragma unroll
                for (int i = 0; i < NPF; i++) {
                    const int p = wave + 4 * i;
                    if (p >= np) continue;   // wave-uniform
                    const int ul = 2 * p + (lq >> 1);
                    const bool wr = wr_lane && ul < nu;
                    const VjpUnit& va = vj[uidx_a[i]];
                    const VjpUnit& vb = vj[uidx_b[i]];
                    // the T2' sum b and the fieldmap sum f of this kind, when it gives them
                    float br = 0.f, bi = 0.f, fr = 0.f, fi = 0.f;
                    bool has_b = false, has_f = false;
                    if (k == 1) {
                        br = re_a[i];
                        bi = im_a[i];
                        has_b = true;
                    } else if (k == 2) {
                        fr = re_a[i] + re_b[i];
                        fi = im_a[i] + im_b[i];
                        has_f = true;
                        if (b_from_f) {
                            // mode 1: tau = alpha tb0 (no second part); mode 2: tau = alpha tb0 on the
                            // lines with tb0 >= 0 (part a) and -alpha tb0 on the others (part b)
                            br = alpha * (re_a[i] - re_b[i]);
                            bi = alpha * (im_a[i] - im_b[i]);
                            has_b = true;
                        }
                    }
                    if (k == 0) {
                        GAS float* gimg = gst((float*)((lq >> 1) ? vb.gimg : va.gimg));
                        if (wr && gimg) {
                            if constexpr (CPLX) {
                                gimg[2 * (q * P + iv)] = re_a[i];
                                gimg[2 * (q * P + iv) + 1] = im_a[i];
                            } else {
                                gimg[q * P + iv] = re_a[i];
                            }
                        }
                    }
                    if (has_b) {
                        GAS float* gti = gst((lq >> 1) ? vb.gti : va.gti);
                        if (wr && gti) {
                            // Re(conj(m) b) = m_r b_r + m_i b_i
                            const float re = CPLX ? m_r[i][q] * br + m_i[i][q] * bi : m_r[i][q] * br;
                            gti[q * P + iv] = re / (ti_ms[q] * ti_ms[q]);
                        }
                    }
                    if (has_f) {
                        // Im(conj(m) f) = m_r f_i - m_i f_r
                        fm[i][0] = __builtin_fmaf(m_r[i][q], fi, fm[i][0]);
                        if constexpr (CPLX) fm[i][1] = __builtin_fmaf(-m_i[i][q], fr, fm[i][1]);
                    }
                }
            }
        }
        // the fieldmap gradient, and zeros where a unit asks for a sum its group does not have
        if (wr_lane) {
#pragma unroll
            for (int i = 0; i < NPF; i++) {
                const int p = wave + 4 * i, ul = 2 * p + (lq >> 1);
                if (p >= np || ul >= nu) continue;
                const VjpUnit& vv = vj[(lq >> 1) ? uidx_b[i] : uidx_a[i]];
                if (vv.gfmap) {
                    const float f = CPLX ? fm[i][0] + fm[i][CM - 1] : fm[i][0];
                    gst(vv.gfmap)[iv] = need2 ? (float)(kTau * (double)f) : 0.0f;
                }
                if (vv.gti && !need1) {
#pragma unroll
                    for (int q = 0; q < Q; q++) gst(vv.gti)[q * P + iv] = 0.0f;
                }
            }
        }
    }
}

template <int Q, bool CPLX>
static hipError_t alaunch_npf(int npf, int nwork, const DevUnit* u, const DevSlice* s, const DevGroup* g,
                              const VjpUnit* v, const WorkItem* w, const int32_t* sets, hipStream_t st) {
    const dim3 grid(nwork);
    switch (npf) {
    case 1: hipLaunchKernelGGL((yadjoint_basis<Q, CPLX, 1>), grid, dim3(256), 0, st, u, s, g, v, w, sets); break;
    case 2: hipLaunchKernelGGL((yadjoint_basis<Q, CPLX, 2>), grid, dim3(256), 0, st, u, s, g, v, w, sets); break;
    case 3: hipLaunchKernelGGL((yadjoint_basis<Q, CPLX, 3>), grid, dim3(256), 0, st, u, s, g, v, w, sets); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// npf = unit pairs per wave (a set holds up to 8 * npf units); Q = 1..4 compartments.  `units` and
// `v` are the adjoint's descriptor arrays; the sets list indices into them.
hipError_t launch_yadjoint_basis(int Q, bool cplx, int npf, int nwork, const DevUnit* u, const DevSlice* s,
                                 const DevGroup* g, const VjpUnit* v, const WorkItem* w, const int32_t* sets,
                                 hipStream_t st) {
    if (nwork <= 0) return hipSuccess;
    switch (Q) {
    case 1: return cplx ? alaunch_npf<1, true>(npf, nwork, u, s, g, v, w, sets, st)
                        : alaunch_npf<1, false>(npf, nwork, u, s, g, v, w, sets, st);
    case 2: return cplx ? alaunch_npf<2, true>(npf, nwork, u, s, g, v, w, sets, st)
                        : alaunch_npf<2, false>(npf, nwork, u, s, g, v, w, sets, st);
    case 3: return cplx ? alaunch_npf<3, true>(npf, nwork, u, s, g, v, w, sets, st)
                        : alaunch_npf<3, false>(npf, nwork, u, s, g, v, w, sets, st);
    case 4: return cplx ? alaunch_npf<4, true>(npf, nwork, u, s, g, v, w, sets, st)
                        : alaunch_npf<4, false>(npf, nwork, u, s, g, v, w, sets, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace mrs
