// Device arithmetic shared by the phasor build and the contraction kernel
// that generates its operands in place: both must produce the same bits.
#pragma once
#include "hv_internal.h"

namespace hv {

// --------------------------------------------------------------- helpers
// fmin/fmax return the non-NaN operand, so a NaN coordinate clamps to 0.
template <typename R>
__device__ inline R clampr(R v, R lo, R hi) { return fmin(fmax(v, lo), hi); }

// Clamp-to-edge bilinear interpolation of beam[y][x] (P x P) at direction
// cosines (x, y).  The texel coordinate and its integer part are formed in
// FP64 and only the fraction is rounded to R (the reference forms the
// coordinate in double and stores it in DTYPE, vis.py:63,69; keeping it in
// FP64 removes the (P-1) 2^-24 quantisation of the interpolation weight).
template <typename R>
__device__ inline R beam_bilinear(const R* __restrict__ b, int P, double x, double y) {
    const double bx = (double)(P - 1) * (0.5 * x + 0.5);
    const double by = (double)(P - 1) * (0.5 * y + 0.5);
    const double px = floor(bx), py = floor(by);
    const R fx = (R)(bx - px), fy = (R)(by - py);
    const double pm = (double)(P - 1);
    const int ix0 = (int)clampr(px, 0.0, pm), ix1 = (int)clampr(px + 1.0, 0.0, pm);
    const int iy0 = (int)clampr(py, 0.0, pm), iy1 = (int)clampr(py + 1.0, 0.0, pm);
    const R v00 = b[iy0 * P + ix0], v01 = b[iy0 * P + ix1];
    const R v10 = b[iy1 * P + ix0], v11 = b[iy1 * P + ix1];
    const R top = fma(fx, v01 - v00, v00);
    const R bot = fma(fx, v11 - v10, v10);
    return fma(fy, bot - top, top);
}

// sin and cos of a phase given in quarter turns t4 (FP32 result, FP64
// reduction): t4 = q + u with q = rint(t4) and |u| <= 1/2 exact in FP64;
// sin/cos(pi/2 u) come from odd/even polynomials in u fitted here by
// iteratively reweighted least squares (Lawson's minimax iteration) on
// [-1/2, 1/2] (max error 8.9e-8 / 7.4e-8 when evaluated in FP32), and the
// quadrant q mod 4 rotates the pair.  q mod 4 is read from the low mantissa
// bits of q + 1.5 * 2^52, exact for |q| < 2^51.
__device__ inline void sincos_quarter_turns(double t4, float* s, float* c) {
    const double q = rint(t4);
    const float u = (float)(t4 - q);
    const int k = (int)(__double2loint(q + 6755399441055744.0) & 3);
    const float u2 = u * u;
    const float ps = u * fmaf(u2, fmaf(u2, fmaf(u2, -0.0046021766029298306f, 0.07968022674322128f),
                                      -0.6459634900093079f), 1.5707963705062866f);
    const float pc = fmaf(u2, fmaf(u2, fmaf(u2, fmaf(u2, 0.0008997500990517437f, -0.020859001204371452f),
                                            0.2536691725254059f), -1.2337005138397217f), 1.0f);
    const bool swap = k & 1;
    const float sv = swap ? pc : ps;
    const float cv = swap ? ps : pc;
    // quadrants 2, 3 negate sin; quadrants 1, 2 negate cos
    *s = __int_as_float(__float_as_int(sv) ^ ((k & 2) << 30));
    *c = __int_as_float(__float_as_int(cv) ^ (((k + 1) & 2) << 30));
}
__device__ inline void sincos_quarter_turns(double t4, double* s, double* c) {
    const double r = t4 - rint(t4);            // [-1/2, 1/2] quarter turns
    sincospi(0.5 * r, s, c);                   // angle pi/2 r, then rotate by the quadrant
    const int k = (int)(__double2loint(rint(t4) + 6755399441055744.0) & 3);
    const double sv = (k & 1) ? *c : *s;
    const double cv = (k & 1) ? *s : *c;
    *s = (k & 2) ? -sv : sv;
    *c = ((k + 1) & 2) ? -cv : cv;
}

// sin and cos of 2 pi t for a phase t in turns (split-mode operands): the
// fraction r = t - floor(t) in [0, 1) is exact in FP64 (one v_fract_f64) and
// rounded once to FP32 (error <= 2^-25 turns), then the hardware sine and
// cosine, which take revolutions (measured max error 1.25e-7 against FP64
// sincospi; < 4e-7 in all, far below the 2^-17 of a hi+lo split).  Three
// instructions replace the polynomials and the quadrant rotation.
__device__ inline void sincos_turns_hw(double t, float* s, float* c) {
    const float r = (float)__builtin_amdgcn_fract(t);
    *s = __builtin_amdgcn_sinf(r);
    *c = __builtin_amdgcn_cosf(r);
}

// Delay phase in turns of a source at (x, y, z) for an antenna whose position
// is already in wavelengths (metres times 1/lambda, rounded once): the dot
// product in FP64.  Every split-mode operand builder forms the phase this
// way, so their operands agree bit for bit.
__device__ inline double phase_turns(double axl, double ayl, double azl, double x, double y, double z) {
    return fma(axl, x, fma(ayl, y, azl * z));
}

// Round-to-nearest-even float -> bf16 / fp16 bit patterns and back.
__device__ inline uint16_t to_bf16(float x) {
    __bf16 h = (__bf16)x;
    return __builtin_bit_cast(uint16_t, h);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ inline uint16_t to_f16(float x) {
    _Float16 h = (_Float16)x;
    return __builtin_bit_cast(uint16_t, h);
}
__device__ inline float from_f16(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }


// hi and lo words of one complex value: the 16-bit (x, y) pairs of v = hi +
// lo, hi = round-to-nearest-even(v) and lo = round-to-nearest-even(v - hi),
// x in the low half.  Packed conversions (one instruction per pair).
typedef float hv_f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 hv_bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 hv_f16x2 __attribute__((ext_vector_type(2)));

template <int MODE>
__device__ inline void split_pair(float re, float im, uint32_t& hi, uint32_t& lo) {
    if constexpr (MODE == MODE_SPLIT_BF16) {
        hi = __builtin_bit_cast(uint32_t, __builtin_convertvector((hv_f32x2){re, im}, hv_bf16x2));
        const float br = __uint_as_float(hi << 16), bm = __uint_as_float(hi & 0xffff0000u);
        lo = __builtin_bit_cast(uint32_t, __builtin_convertvector((hv_f32x2){re - br, im - bm}, hv_bf16x2));
    } else {
        const hv_f16x2 h = __builtin_convertvector((hv_f32x2){re, im}, hv_f16x2);
        hi = __builtin_bit_cast(uint32_t, h);
        const hv_f32x2 b = __builtin_convertvector(h, hv_f32x2);
        lo = __builtin_bit_cast(uint32_t, __builtin_convertvector((hv_f32x2){re - b[0], im - b[1]}, hv_f16x2));
    }
}

// hi and lo words of the phasor am (cs, sn), fixed arithmetic (no
// contraction): hi = RNE(round(am cs), round(am sn)) and lo =
// RNE(fma(am, cs, -hi_x), fma(am, sn, -hi_y)), so every kernel that builds
// operands produces the same bits.
template <int MODE>
__device__ inline void split_phasor(float am, float cs, float sn, uint32_t& hi, uint32_t& lo) {
#pragma clang fp contract(off)
    const float re = am * cs, im = am * sn;
    float br, bm;
    if constexpr (MODE == MODE_SPLIT_BF16) {
        hi = __builtin_bit_cast(uint32_t, __builtin_convertvector((hv_f32x2){re, im}, hv_bf16x2));
        br = __uint_as_float(hi << 16);
        bm = __uint_as_float(hi & 0xffff0000u);
        lo = __builtin_bit_cast(uint32_t,
                                __builtin_convertvector((hv_f32x2){fmaf(am, cs, -br), fmaf(am, sn, -bm)}, hv_bf16x2));
    } else {
        const hv_f16x2 h = __builtin_convertvector((hv_f32x2){re, im}, hv_f16x2);
        hi = __builtin_bit_cast(uint32_t, h);
        const hv_f32x2 b = __builtin_convertvector(h, hv_f32x2);
        lo = __builtin_bit_cast(uint32_t,
                                __builtin_convertvector((hv_f32x2){fmaf(am, cs, -b[0]), fmaf(am, sn, -b[1])}, hv_f16x2));
    }
}

// Split 4 complex values into a 32-byte group [8 hi | 8 lo] of 16-bit halves.
template <int MODE>
__device__ inline void split_group(const float* re, const float* im, uint4* out) {
    uint32_t h[4], l[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) split_pair<MODE>(re[j], im[j], h[j], l[j]);
    out[0] = make_uint4(h[0], h[1], h[2], h[3]);
    out[1] = make_uint4(l[0], l[1], l[2], l[3]);
}

}  // namespace hv
