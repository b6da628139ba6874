// devcommon.h -- device helpers shared by the y-stage kernels: vector types, global address-space
// casts, the split-BF16 operand split and three-term MFMA product, the FP64-reduced sincos of a
// phase in turns, direct-to-LDS copies and partial memory waits.  Not part of the public ABI.
#pragma once
#include "internal.h"

namespace mrs {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;
typedef __attribute__((ext_vector_type(2))) uint32_t u32x2;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(2))) double f64x2;

// Pointers read from device descriptors are generic; casting them to the global address space
// lets the compiler emit global loads with a scalar base instead of flat loads.
#define GAS __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const GAS T* gld(const T* p) { return (const GAS T*)p; }
template <class T>
__device__ __forceinline__ GAS T* gst(T* p) { return (GAS T*)p; }
// Element `i` of a device array addressed as uniform base + 32-bit byte offset (the offset is formed
// in 32-bit arithmetic, so the load takes the saddr form: no per-lane 64-bit address arithmetic).
// Only for arrays below 4 GiB, which every per-slice plane and unit image is.
template <class V, class T>
__device__ __forceinline__ const GAS V* at32(const T* base, uint32_t i) {
    return (const GAS V*)((const GAS char*)base + i * (uint32_t)sizeof(T));
}

static constexpr double kTau = 6.283185307179586476925286766559;
static constexpr float kLog2e = 1.4426950408889634f;

// ----------------------------------------------------------------------------- small helpers

__device__ __forceinline__ uint32_t bf16_bits(float v) {
    __bf16 h = (__bf16)v;  // round to nearest even (v_cvt_pk_bf16_f32)
    return (uint32_t)__builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float bf16_float(uint32_t b) { return __builtin_bit_cast(float, b << 16); }

// (a, b) -> packed hi = bf16(a) | bf16(b) << 16 (one v_cvt_pk_bf16_f32, round to nearest even),
// lo = the same of the residuals a - hi_a, b - hi_b.
__device__ __forceinline__ void split2(float a, float b, uint32_t& hi, uint32_t& lo) {
    const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
    const float ha = __builtin_bit_cast(float, h << 16), hb = __builtin_bit_cast(float, h & 0xffff0000u);
    hi = h;
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a - ha, b - hb}, bf16x2));
}

// One complex value of the split-precision G buffer: {hi, lo} of split2(re, im).
__device__ __forceinline__ u32x2 split_g(float re, float im) {
    uint32_t h, l;
    split2(re, im, h, l);
    return u32x2{h, l};
}

// A-operand rows of a complex contraction out = sum S * v, split into (re row, im row):
//   re row: [S_re, -S_im] per complex k;  im row: [S_im, S_re].
// `p` holds four packed (re | im<<16) complex values.
__device__ __forceinline__ u32x4 rows_of(u32x4 p, int part) {
    u32x4 r;
    if (part == 0) {
#pragma unroll
        for (int i = 0; i < 4; i++) r[i] = p[i] ^ 0x80000000u;
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++) r[i] = (p[i] >> 16) | (p[i] << 16);
    }
    return r;
}

// One bf16 MFMA 16x16x32 with FP32 accumulation of packed operands.
__device__ __forceinline__ f32x4 mfma_bf16(u32x4 a, u32x4 b, f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                   0, 0, 0);
}

__device__ __forceinline__ f32x4 mfma3(u32x4 ahi, u32x4 alo, u32x4 bhi, u32x4 blo, f32x4 acc) {
    bf16x8 Ah = __builtin_bit_cast(bf16x8, ahi), Al = __builtin_bit_cast(bf16x8, alo);
    bf16x8 Bh = __builtin_bit_cast(bf16x8, bhi), Bl = __builtin_bit_cast(bf16x8, blo);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al, Bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah, Bh, acc, 0, 0, 0);
    return acc;
}

// sin and cos of 2*pi*cyc, two evaluations of the FP64-reduced argument:
//  poly: reduce in FP64 to the nearest quarter turn; the FP32 residual f in [-1/8, 1/8] turn is
//        evaluated by Taylor polynomials of sin and cos about 0 (degree 9 and 10 in theta = 2 pi f,
//        truncation below 2e-9 at |theta| = pi/4), then rotated by the quarter turns.
//  hw:   reduce in FP64 to [-1/2, 1/2] turn and use v_sin_f32 / v_cos_f32, which take turns.
__device__ __forceinline__ void sincos_turns_hw(double cyc, float& s, float& c) {
    const float r = (float)(cyc - __builtin_rint(cyc));
    s = __builtin_amdgcn_sinf(r);
    c = __builtin_amdgcn_cosf(r);
}

__device__ __forceinline__ void sincos_turns_poly(double cyc, float& s, float& c) {
    const double q4 = __builtin_rint(cyc * 4.0);
    const float f = (float)__builtin_fma(q4, -0.25, cyc);
    const int k = (int)q4 & 3;
    const float th = f * 6.2831853071795864769f;
    const float t2 = th * th;
    float ps = __builtin_fmaf(t2, 2.7557319223985893e-06f, -1.9841269841269841e-04f);
    ps = __builtin_fmaf(t2, ps, 8.3333333333333333e-03f);
    ps = __builtin_fmaf(t2, ps, -1.6666666666666667e-01f);
    ps = t2 * ps;
    const float sn = __builtin_fmaf(th, ps, th);
    float pc = __builtin_fmaf(t2, -2.7557319223985891e-07f, 2.4801587301587302e-05f);
    pc = __builtin_fmaf(t2, pc, -1.3888888888888889e-03f);
    pc = __builtin_fmaf(t2, pc, 4.1666666666666667e-02f);
    pc = __builtin_fmaf(t2, pc, -0.5f);
    const float cs = __builtin_fmaf(t2, pc, 1.0f);
    const float s1 = (k & 1) ? cs : sn;
    const float c1 = (k & 1) ? sn : cs;
    s = (k & 2) ? -s1 : s1;
    c = ((k + 1) & 2) ? -c1 : c1;
}

// v_sin/v_cos on the FP64-reduced argument: max abs error 1.7e-7 measured on MI355X against FP64
// over 2e5 arguments in [-300, 300] turns (the polynomial: 9.4e-8), far inside the 1e-5 budget.
#ifndef MRS_HW_SINCOS
#define MRS_HW_SINCOS 1
#endif
__device__ __forceinline__ void sincos_turns(double cyc, float& s, float& c) {
#if MRS_HW_SINCOS
    sincos_turns_hw(cyc, s, c);
#else
    sincos_turns_poly(cyc, s, c);
#endif
}

// Direct-to-LDS copy of `bytes` (4 or 16) per lane from uniform base + per-lane byte offset into
// lds + lane*bytes.
template <int BYTES>
__device__ __forceinline__ void lds_copy(const void* base, uint32_t off, uint32_t* lds) {
    static_assert(BYTES == 4 || BYTES == 16, "dword or dwordx4 copies");
    const GAS void* src = (const GAS void*)((const GAS char*)base + off);
    __attribute__((address_space(3))) void* dst = (__attribute__((address_space(3))) void*)lds;
    if constexpr (BYTES == 16)
        __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
    else
        __builtin_amdgcn_global_load_lds(src, dst, 4, 0, 0);
}

// s_waitcnt vmcnt(N) with the other counters left open (gfx9 encoding).
template <int N>
__device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x70 | 0xF00);
}

}  // namespace mrs
