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

}  // namespace mrs
