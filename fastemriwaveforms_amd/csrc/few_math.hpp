// Small host+device math helpers of the mode sums.
//
// sincos_cw(x): sin and cos of x in radians with a Cody-Waite reduction
// (|x| up to ~1e8; the mode phases reach 1e5-1e7 rad).
// sincos_turns(f): sin and cos of 2*pi*f for f in [-1/2, 1/2] (an angle in
// turns already reduced in FP64).  The quadrant q = rint(4f) is removed
// exactly (f - q/4 is exact in binary floating point), leaving
// |2*pi*r| <= pi/4, where the Taylor series of sin and cos are summed by
// Horner's rule in FP64 up to x^17 and x^18 (truncation error below
// (pi/4)^19/19! ~ 1e-19), then rotated back by the quadrant.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace few {

// sin and cos of x in [-pi/4, pi/4] (radians), rotated by q quarter turns:
// Taylor series summed by Horner's rule up to x^17 and x^18 (truncation
// below (pi/4)^19/19! ~ 1e-19).
__host__ __device__ inline void sincos_poly(double x, int q, double &s, double &c)
{
    const double x2 = x * x;
    // sin x = x (1 - x^2/3! + x^4/5! - ... + x^16/17!)
    double ps = 2.8114572543455207632e-15;            // 1/17!
    ps = ps * x2 - 7.6471637318198164759e-13;         // -1/15!
    ps = ps * x2 + 1.6059043836821614599e-10;         // 1/13!
    ps = ps * x2 - 2.5052108385441718775e-08;         // -1/11!
    ps = ps * x2 + 2.7557319223985890653e-06;         // 1/9!
    ps = ps * x2 - 1.9841269841269841270e-04;         // -1/7!
    ps = ps * x2 + 8.3333333333333333333e-03;         // 1/5!
    ps = ps * x2 - 1.6666666666666666667e-01;         // -1/3!
    const double sn = x + x * x2 * ps;
    // cos x = 1 - x^2/2! + x^4/4! - ... + x^18/18!
    double pc = 1.5619206968586225351e-16;            // 1/18!
    pc = pc * x2 - 4.7794773323873852974e-14;         // -1/16!
    pc = pc * x2 + 1.1470745597729724714e-11;         // 1/14!
    pc = pc * x2 - 2.0876756987868098979e-09;         // -1/12!
    pc = pc * x2 + 2.7557319223985890653e-07;         // 1/10!
    pc = pc * x2 - 2.4801587301587301587e-05;         // -1/8!
    pc = pc * x2 + 1.3888888888888888889e-03;         // 1/6!
    pc = pc * x2 - 4.1666666666666666667e-02;         // -1/4!
    pc = pc * x2 + 0.5;
    const double cs = 1.0 - x2 * pc;
    // rotate by q quarter turns (branch-free): odd q swaps sin and cos,
    // q = 2, 3 negate sin, q = 1, 2 negate cos
    const int qi = q & 3;
    const bool swap = qi & 1;
    const double s0 = swap ? cs : sn;
    const double c0 = swap ? sn : cs;
    s = (qi & 2) ? -s0 : s0;
    c = ((qi + 1) & 2) ? -c0 : c0;
}

// Shorter variant for the mixed-precision (FP16-split) paths, whose values
// carry ~1e-7 relative error: Taylor series to x^11 and x^12 (truncation below
// (pi/4)^13/13! ~ 7e-12), same quadrant handling.
__host__ __device__ inline void sincos_poly_short(double x, int q, double &s, double &c)
{
    const double x2 = x * x;
    double ps = -2.5052108385441718775e-08;           // -1/11!
    ps = ps * x2 + 2.7557319223985890653e-06;         // 1/9!
    ps = ps * x2 - 1.9841269841269841270e-04;         // -1/7!
    ps = ps * x2 + 8.3333333333333333333e-03;         // 1/5!
    ps = ps * x2 - 1.6666666666666666667e-01;         // -1/3!
    const double sn = x + x * x2 * ps;
    double pc = -2.0876756987868098979e-09;           // -1/12!
    pc = pc * x2 + 2.7557319223985890653e-07;         // 1/10!
    pc = pc * x2 - 2.4801587301587301587e-05;         // -1/8!
    pc = pc * x2 + 1.3888888888888888889e-03;         // 1/6!
    pc = pc * x2 - 4.1666666666666666667e-02;         // -1/4!
    pc = pc * x2 + 0.5;
    const double cs = 1.0 - x2 * pc;
    const int qi = q & 3;
    const bool swap = qi & 1;
    const double s0 = swap ? cs : sn;
    const double c0 = swap ? sn : cs;
    s = (qi & 2) ? -s0 : s0;
    c = ((qi + 1) & 2) ? -c0 : c0;
}

__host__ __device__ inline void sincos_turns_short(double f, double &s, double &c)
{
    const double q = rint(4.0 * f);
    const double r = f - 0.25 * q;
    sincos_poly_short(r * 6.283185307179586476925286766559, (int)q, s, c);
}

// One fused multiply-add a * b + c (rounded once, as fma()) whose addend is a
// wave-uniform constant.  On the device it is pinned to the three-address
// v_fma_f64 with the addend read from an SGPR pair: the compiler otherwise
// copies polynomial coefficients into VGPRs and spends a move on each
// two-address v_fmac_f64.
__host__ __device__ __forceinline__ double fma_cs(double a, double b, double c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
#else
    return fma(a, b, c);
#endif
}

// sin and cos of 2*pi*f for f in [-1/2, 1/2] (turns already reduced) without
// a quadrant reduction, for the FP16-split paths: x = 2*pi*f in [-pi, pi],
// sin x = x S(x^2) and cos x = C(x^2) with S and C of degree 8 in x^2 -- a
// Chebyshev least-squares fit on [0, pi^2] (sin(x)/x weighted by sqrt(x)),
// converted to monomials.  Absolute error below 3.1e-13 (sin) and 2.6e-12
// (cos) over the whole range, against 7e-12 truncation for
// sincos_turns_short, in 21 operations instead of about 35 (no quadrant,
// swap or sign selection).
__host__ __device__ inline void sincos_turns_circle(double f, double &s, double &c)
{
    const double x = f * 6.283185307179586476925286766559;
    const double u = x * x;
    double ps = 2.459350323376714e-15;
    ps = fma_cs(ps, u, -7.581087442893419e-13);
    ps = fma_cs(ps, u, 1.6052202305314846e-10);
    ps = fma_cs(ps, u, -2.505168645272782e-08);
    ps = fma_cs(ps, u, 2.7557303499119424e-06);
    ps = fma_cs(ps, u, -0.00019841269497515153);
    ps = fma_cs(ps, u, 0.008333333329272535);
    ps = fma_cs(ps, u, -0.16666666666449323);
    ps = fma_cs(ps, u, 0.9999999999996599);
    s = x * ps;
    double pc = 4.131222830361208e-14;
    pc = fma_cs(pc, u, -1.1353068040299952e-11);
    pc = fma_cs(pc, u, 2.0865037866218456e-09);
    pc = fma_cs(pc, u, -2.7556630082695456e-07);
    pc = fma_cs(pc, u, 2.4801563123496047e-05);
    pc = fma_cs(pc, u, -0.0013888888401028148);
    pc = fma_cs(pc, u, 0.041666666615207106);
    pc = fma_cs(pc, u, -0.4999999999769597);
    c = fma_cs(pc, u, 0.9999999999974767);
}

__host__ __device__ inline void sincos_turns(double f, double &s, double &c)
{
    const double q = rint(4.0 * f);
    const double r = f - 0.25 * q;
    sincos_poly(r * 6.283185307179586476925286766559, (int)q, s, c);  // 2*pi (double)
}

// sin and cos of x (radians) for |x| <= 2^26 pi/2 (~1.05e8): Cody-Waite
// reduction x = k pi/2 + r with pi/2 = P1 + P2 + P3 (26, 26 and 53
// significant bits, from pi to 80 digits; |pi/2 - P1 - P2 - P3| < 2e-33), so
// k P1 and k P2 are exact for |k| <= 2^26 and the fused steps lose nothing;
// then sincos_poly.  Returns false (outputs untouched) beyond that range.
__host__ __device__ inline bool sincos_cw(double x, double &s, double &c)
{
    const double k = rint(x * 0.63661977236758134308);  // 2/pi
    if (!(fabs(k) <= 67108864.0)) return false;          // 2^26 (also rejects NaN)
    double r = fma(-k, 1.5707963109016418457, x);        // P1 = 0x1.921fb5p+0
    r = fma(-k, 1.5893254712295857e-08, r);              // P2 = 0x1.110b46p-26
    r = fma(-k, 6.123233995736766e-17, r);               // P3 = 0x1.1a62633145c07p-54
    sincos_poly(r, (int)(long long)k, s, c);
    return true;
}

// Angle in turns reduced to [-1/2, 1/2]: a - rint(a).
__host__ __device__ inline double reduce_turns(double a) { return a - rint(a); }

// A float the optimiser cannot trace back to the double it was rounded from.
// Without it, (_Float16)(float)d is folded into one double -> half rounding,
// which gfx950 has no instruction for (~30 integer instructions each); with
// it, the conversions are the hardware float -> half ones (round to nearest).
// The asm statement is empty: it emits no instruction.
__device__ __forceinline__ float f32_opaque(float x)
{
    asm volatile("" : "+v"(x));
    return x;
}

typedef float v2f __attribute__((ext_vector_type(2)));
typedef _Float16 v2h __attribute__((ext_vector_type(2)));

// Low FP16 parts of a pair x split as h = fp16(x): fp16(x - h) per lane,
// computed by v_fma_mix{lo,hi}_f16 (x - h is exact in FP32, so the one
// rounding is the FP16 one, as in fp16((float)x - (float)h)); two
// instructions instead of two conversions back, two subtractions and a pack.
__device__ __forceinline__ v2h f16_split_lo(v2h h, v2f x)
{
    v2h r;
    asm("v_fma_mixlo_f16 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]\n\t"
        "v_fma_mixhi_f16 %0, %1, -1.0, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
        : "=&v"(r)
        : "v"(h), "v"(x.x), "v"(x.y));
    return r;
}

}  // namespace few
