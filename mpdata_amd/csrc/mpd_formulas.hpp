// Device formulas of the MPDATA step, written in the reference's evaluation order so
// that, compiled with -ffp-contract=off and IEEE division, every kernel that uses them
// reproduces the reference's FP64 rounding bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>

namespace mpd {

// smolar_int.h:25 -- guard added to every ratio denominator.
constexpr double MINV = DBL_MIN;

// smolar_int.h:8 -- (a > 0 ? a : -a).
__device__ __forceinline__ double ab(double a) { return a > 0 ? a : -a; }

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

// uu.c:18-32 (vv.c:21-35, ww.c:180-194) -- antidiffusive pseudo-velocity of one face:
// a = previous-pass Courant number of the face, p/m = the cell values above/below the
// face, s1/s2 = averaged transverse velocities, n/d = transverse gradient ratios.
__device__ __forceinline__ double antidiff(double div, double a, double p, double m,
                                           double s1, double n1, double d1, double s2,
                                           double n2, double d2)
{
    double r = (ab(a) * (1.0 - ab(a))) * (p - m) / (p + m + MINV);
    return (-div + r) - 0.5 * (a * (s1 * n1 / (d1 + MINV) + (s2 * n2 / (d2 + MINV))));
}

__device__ __forceinline__ double pos(double a) { return a > 0.0 ? a : 0.0; }
__device__ __forceinline__ double neg(double a) { return a < 0.0 ? a : 0.0; }
__device__ __forceinline__ double mx2(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double mn2(double a, double b) { return a < b ? a : b; }

// Nonoscillatory limiter, Smolarkiewicz & Grabowski (1990): limited face velocity
// from the bounds ratios of the cells below (l) and above (r) the face.
__device__ __forceinline__ double limit_face(double u, double bup_l, double bdn_l,
                                             double bup_r, double bdn_r)
{
    return mn2(1.0, mn2(bdn_l, bup_r)) * pos(u) + mn2(1.0, mn2(bup_l, bdn_r)) * neg(u);
}

}  // namespace mpd
