// Device formulas of the MPDATA step, written in the reference's evaluation order.
//
// Two arithmetic flavours share this text (the including file picks one by defining
// MPD_ARITH before inclusion):
//   exact  compiled -ffp-contract=off with IEEE division: every kernel reproduces the
//          reference's FP64 rounding bit for bit (up to the sign of zero results, which
//          never changes a non-zero value: every denominator carries +MINV > 0).
//   fast   compiled -ffp-contract=fast with a Newton-refined hardware reciprocal for
//          division: a few ulp per operation, held to <= 1e-12 normwise per step
//          against the oracle by the tests.
// Selections (|a|, max, min, positive/negative part) use the hardware min/max/abs
// instead of compare-and-select: identical values except the sign of zero.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>

#ifndef MPD_ARITH
#define MPD_ARITH exact
#endif

namespace mpd {
namespace MPD_ARITH {

// smolar_int.h:25 -- guard added to every ratio denominator.
constexpr double MINV = DBL_MIN;

// smolar_int.h:8 -- (a > 0 ? a : -a); fabs differs only for a = +0.
__device__ __forceinline__ double ab(double a) { return __builtin_fabs(a); }

#ifdef MPD_FAST_DIV
// n / d for d >= DBL_MIN (every denominator here is a sum of non-negative cell values
// plus MINV, or a positive flux sum plus MINV): hardware reciprocal estimate r of
// relative error e0 (v_rcp_f64: about 2^-26 on gfx950), the quotient q = n r, then one
// residual correction q <- q + r(n - d q), whose error is e0^2 plus the rounding of the
// two fused steps.  Measured on 4M quotients spanning the kernels' magnitudes: at most
// 18 ulp, 0.6 ulp on average, of the IEEE quotient (a Newton step on r first would make
// it exact there, for two more instructions per division).
__device__ __forceinline__ double dv(double n, double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double q = n * r;
    return __builtin_fma(r, __builtin_fma(-d, q, n), q);
}
#else
__device__ __forceinline__ double dv(double n, double d) { return n / d; }
#endif

// smolar_pass.c:104 -- twice the donor flux through a face with Courant number v from
// the lower cell value l and the upper cell value r.
__device__ __forceinline__ double F2(double l, double r, double v)
{
    double av = ab(v);
    return l * (v + av) + r * (v - av);
}

// smolar_pass.c:106-118 -- donor update from the six face fluxes (x lower/upper, ...).
__device__ __forceinline__ double donor(double c, double cxl, double cxr, double cyl,
                                        double cyr, double czl, double czr, double uxl,
                                        double uxr, double vyl, double vyr, double wzl,
                                        double wzr)
{
    double fx = F2(c, cxr, uxr) - F2(cxl, c, uxl);
    double fy = F2(c, cyr, vyr) - F2(cyl, c, vyl);
    double fz = F2(c, czr, wzr) - F2(czl, c, wzl);
    return c - 0.5 * ((fx + fy) + fz);
}

// donor() with the lower z face's doubled flux F2(czl, c, wzl) given: a streaming kernel
// carries it from the plane below, where it was that plane's upper z face, F2(c, czr,
// wzr), which this returns in fzu2 (the same operations, so the same values).
__device__ __forceinline__ double donor_zc(double c, double cxl, double cxr, double cyl,
                                           double cyr, double czr, double uxl, double uxr,
                                           double vyl, double vyr, double wzr, double fzl2,
                                           double& fzu2)
{
    double fx = F2(c, cxr, uxr) - F2(cxl, c, uxl);
    double fy = F2(c, cyr, vyr) - F2(cyl, c, vyl);
    fzu2 = F2(c, czr, wzr);
    double fz = fzu2 - fzl2;
    return c - 0.5 * ((fx + fy) + fz);
}

// uu.c:18-32 (vv.c:21-35, ww.c:180-194) -- antidiffusive pseudo-velocity of one face:
// a = previous-pass Courant number of the face, p/m = the cell values above/below the
// face, s1/s2 = averaged transverse velocities, n/d = transverse gradient ratios.
__device__ __forceinline__ double antidiff(double div, double a, double p, double m,
                                           double s1, double n1, double d1, double s2,
                                           double n2, double d2)
{
    double r = dv((ab(a) * (1.0 - ab(a))) * (p - m), p + m + MINV);
    return (-div + r) - 0.5 * (a * (dv(s1 * n1, d1 + MINV) + dv(s2 * n2, d2 + MINV)));
}

#ifdef MPD_FAST_DIV
// antidiff() with its constant factors gathered (fast arithmetic): from the unscaled
// divergence sum D (div = a D / 4) and the unscaled transverse sums X, Y (s1 = X / 4,
// s2 = Y / 4) it forms r - a (2 D + X n1 / d1 + Y n2 / d2) / 8, the same expression: the
// factors are powers of two (exact scalings), so only the roundings of the fused steps
// differ, and five operations of each face go (|a| (1 - |a|) is one fused step,
// |a| - |a| |a|).
__device__ __forceinline__ double antidiff_s(double D, double a, double p, double m, double X,
                                             double n1, double d1, double Y, double n2,
                                             double d2)
{
    double r = dv(__builtin_fma(-ab(a), ab(a), ab(a)) * (p - m), p + m + MINV);
    double s = dv(X * n1, d1 + MINV) + dv(Y * n2, d2 + MINV);
    return __builtin_fma(-0.125 * a, __builtin_fma(2.0, D, s), r);
}
#endif

__device__ __forceinline__ double pos(double a) { return __builtin_fmax(a, 0.0); }
__device__ __forceinline__ double neg(double a) { return __builtin_fmin(a, 0.0); }
__device__ __forceinline__ double mx2(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double mn2(double a, double b) { return __builtin_fmin(a, b); }

// Nonoscillatory limiter, Smolarkiewicz & Grabowski (1990): limited face velocity
// from the bounds ratios of the cells below (l) and above (r) the face.
__device__ __forceinline__ double limit_face(double u, double bup_l, double bdn_l,
                                             double bup_r, double bdn_r)
{
    return mn2(1.0, mn2(bdn_l, bup_r)) * pos(u) + mn2(1.0, mn2(bup_l, bdn_r)) * neg(u);
}

// The same for ratios already clamped to at most 1 where they were computed (min is
// exact, so min(1, min(a, b)) = min(min(1, a), min(1, b)): the same values, and the clamp
// is paid once per cell instead of four times per face pair).
__device__ __forceinline__ double limit_face_clamped(double u, double bup_l, double bdn_l,
                                                     double bup_r, double bdn_r)
{
    return mn2(bdn_l, bup_r) * pos(u) + mn2(bup_l, bdn_r) * neg(u);
}

// Inflow and outflow sums of one cell from its six donor fluxes (0.5 F2; lower and upper
// face of x, y, z).  The adjoint recomputes them with this same function, so its values
// are those of the forward bit for bit.
__device__ __forceinline__ void flux_sums(double fxl, double fxr, double fyl, double fyr,
                                          double fzl, double fzr, double& ain, double& aout)
{
#ifdef MPD_FAST_DIV
    // fast arithmetic: ain + aout = sum of |f| over the six faces and ain - aout = the net
    // inflow (f_lower - f_upper summed), so both follow from two sums of six (the absolute
    // values are operand modifiers) instead of twelve selections; clamped at zero, where
    // the difference of the rounded sums can fall a rounding below it
    const double sa = ((ab(fxl) + ab(fxr)) + (ab(fyl) + ab(fyr))) + (ab(fzl) + ab(fzr));
    const double sd = ((fxl - fxr) + (fyl - fyr)) + (fzl - fzr);
    ain = pos(0.5 * (sa + sd));
    aout = pos(0.5 * (sa - sd));
#else
    ain = ((pos(fxl) - neg(fxr)) + (pos(fyl) - neg(fyr))) + (pos(fzl) - neg(fzr));
    aout = ((pos(fxr) - neg(fxl)) + (pos(fyr) - neg(fyl))) + (pos(fzr) - neg(fzl));
#endif
}

// Limiter ratios of one cell from the bounds hi/lo, its value c and the six donor
// fluxes of the pass: bup = (hi - c)/(Ain + MINV), bdn = (c - lo)/(Aout + MINV).
__device__ __forceinline__ void beta(double hi, double lo, double c, double fxl, double fxr,
                                     double fyl, double fyr, double fzl, double fzr,
                                     double& bup, double& bdn)
{
    double ain, aout;
    flux_sums(fxl, fxr, fyl, fyr, fzl, fzr, ain, aout);
    bup = dv(hi - c, ain + MINV);
    bdn = dv(c - lo, aout + MINV);
}

// beta() from the doubled donor fluxes F2 of the six faces.  The exact build halves each
// flux first, as beta()'s callers do (the same operations, so the same bits); the fast
// build folds the six halvings into the two sums (a power-of-two scaling, exact: the
// same values as halving first, seven operations fewer).
__device__ __forceinline__ void beta2(double hi, double lo, double c, double Fxl, double Fxr,
                                      double Fyl, double Fyr, double Fzl, double Fzr,
                                      double& bup, double& bdn)
{
#ifdef MPD_FAST_DIV
    const double sa = ((ab(Fxl) + ab(Fxr)) + (ab(Fyl) + ab(Fyr))) + (ab(Fzl) + ab(Fzr));
    const double sd = ((Fxl - Fxr) + (Fyl - Fyr)) + (Fzl - Fzr);
    // pos(x / 4) + MINV as max(x / 4 + MINV, MINV), the scaling and the guard in one
    // fused step (x / 4 is exact, so the sums are those of the unfused form)
    const double dn_in = mx2(__builtin_fma(0.25, sa + sd, MINV), MINV);
    const double dn_out = mx2(__builtin_fma(0.25, sa - sd, MINV), MINV);
    bup = dv(hi - c, dn_in);
    bdn = dv(c - lo, dn_out);
#else
    beta(hi, lo, c, 0.5 * Fxl, 0.5 * Fxr, 0.5 * Fyl, 0.5 * Fyr, 0.5 * Fzl, 0.5 * Fzr, bup, bdn);
#endif
}

}  // namespace MPD_ARITH
}  // namespace mpd
