// Discrete adjoint of the periodic MPDATA step (iord = 2, with or without the
// nonoscillatory limiter) and the checkpointed receptor-loss gradient driver.
// No counterpart in the reference.
//
// One step is  c_out = A(c_in) + q  with fixed face Courant numbers:
//   c1  = donor(c_in; U,V,W)                       first (donor) pass
//   A   = antidiff(c1; U,V,W)                      antidiffusive face velocities
//   bup, bdn = ratios(c_in, c1, A)                 limiter (nonoscillatory option)
//   Al  = limit(A; bup, bdn)
//   c_out = donor(c1; Al) + q                      corrective pass
// The forward intermediates (c1, A, the limiter's ratios, flux sums and arg-max codes)
// come from the fused forward kernels, which store them during the recompute of each
// checkpoint segment.  The adjoint runs the stages backwards, each as a gather over the
// cells or faces that depend on a value (no atomics, deterministic).  Every kink (|v|, max/min,
// positive/negative part, the limiter's min-selection) follows the branch the forward
// took, with the oracle's conventions (oracle/cpu_adj.c), so the GPU adjoint is the
// transpose of the oracle's tangent-linear model.  Arithmetic is unfused IEEE FP64.
#include <algorithm>
#include <memory>
#include <vector>

#include "mpd_formulas.hpp"
#include "mpd_fused.hpp"
#include "mpd_grid.hpp"
#include "mpd_tile.hpp"

namespace mpd {
int ensure_fused_scratch(mpd_grid_s* g);
int periodic_step(mpd_grid_s* g);
}  // namespace mpd

using namespace mpd;
using namespace mpd::exact;

namespace {

constexpr int TPB = 256;
unsigned nblk(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }

__device__ __forceinline__ double sg(double a) { return a > 0 ? 1.0 : -1.0; }

// partials of f = 0.5*F2(l, r, v) with respect to l, r, v
__device__ __forceinline__ void fpart(double l, double r, double v, double& gl, double& gr,
                                      double& gv)
{
    double s = sg(v), av = ab(v);
    gl = 0.5 * (v + av);
    gr = 0.5 * (v - av);
    gv = 0.5 * (l * (1.0 + s) + r * (1.0 - s));
}

// mirrors mn2(1, mn2(a, b)): 0 = the constant 1, 1 = a, 2 = b
__device__ __forceinline__ int pick(double a, double b)
{
    double m = a < b ? a : b;
    int w = a < b ? 1 : 2;
    return 1.0 < m ? 0 : w;
}

// Both sides of the stage-1 adjoint of a z face between a cell t and the cell r above it
// in the same column (the donor gather and the limiter terms of k_adj_lim_t): the terms
// t adds for its upper face and those r adds for its lower face.  They are the same
// face, so the same flux, selection and products; the tiled kernel forms them once per
// face and carries r's share to the next plane of the column.
struct ZFace {
    double f;              // the unlimited donor flux 0.5 F2
    double glf, sdu, suu;  // t: donor-gather term, sd and su terms (selection w = 1)
    double lam;            // lambda of the limited face
    double grlf, sul, sdl; // r: donor-gather term, su and sd terms (selection w = 2)
};

__device__ __forceinline__ ZFace zface1(double ct, double c_r, double a_, double lt,
                                        double lo_r, double bu_t, double bd_t, double bu_r,
                                        double bd_r)
{
    ZFace z;
    z.f = 0.5 * F2(ct, c_r, a_);
    const bool pf = a_ > 0, nf = a_ < 0;
    const double x = pf ? bd_t : bu_t, y = pf ? bu_r : bd_r;
    const int w = pick(x, y);
    const double m = (pf || nf) ? (w == 0 ? 1.0 : (w == 1 ? x : y)) : 0.0;
    const double au = (pf || nf) ? m * a_ : 0.0;
    double gl, gr, gv;
    fpart(ct, c_r, au, gl, gr, gv);
    const double lf = lo_r - lt;
    z.glf = gl * lf;
    z.grlf = gr * lf;
    const double Lf = gv * lf;
    const double sel1 = (w == 1 && (pf || nf)) ? a_ * Lf : 0.0;
    const double sel2 = (w == 2 && (pf || nf)) ? a_ * Lf : 0.0;
    z.sdu = pf ? sel1 : 0.0;
    z.suu = nf ? sel1 : 0.0;
    z.sul = pf ? sel2 : 0.0;
    z.sdl = nf ? sel2 : 0.0;
    z.lam = m * Lf;
    return z;
}

// Stage 2 of k_adj_lim_t on the z face between a cell t and the cell r above it (the
// flux-sum adjoint): t's donor-gather term gl*lf and its face term gv*lf, and the term
// gr*lf that r adds for its lower face -- formed once per face, r's share carried.
__device__ __forceinline__ void zface2(double ct, double c_r, double a_, double li_t,
                                       double lx_t, double li_r, double lx_r, double& glf,
                                       double& gvlf, double& grlf)
{
    // (lx_t [f>0] - li_t [f<0]) + (li_r [f>0] - lx_r [f<0]), the zero terms dropped
    const double f = 0.5 * F2(ct, c_r, a_);
    const double lf = f > 0 ? lx_t + li_r : (f < 0 ? -(li_t + lx_r) : 0.0);
    double gl, gr, gv;
    fpart(ct, c_r, a_, gl, gr, gv);
    glf = gl * lf;
    gvlf = gv * lf;
    grlf = gr * lf;
}

struct Geo {
    int nx, ny, nz;
    size_t np, n;
    __device__ void ijk(size_t t, int& i, int& j, int& k) const
    {
        i = (int)(t % nx);
        j = (int)((t / nx) % ny);
        k = (int)(t / np);
    }
    // wrapped index of a cell at most two cells outside the grid (every stencil here)
    __device__ static int wrap2(int a, int n)
    {
        a += a < 0 ? n : 0;
        a += a < 0 ? n : 0;
        a -= a >= n ? n : 0;
        a -= a >= n ? n : 0;
        return a;
    }
    __device__ size_t at(int i, int j, int k) const
    {
        return (size_t)wrap2(k, nz) * np + (size_t)wrap2(j, ny) * nx + wrap2(i, nx);
    }
};

// The three faces above cell (i,j,k): stencil cells and transverse strokes
// (the interior formulas of uu.c:81-87, vv.c:85-92, ww.c:248-255).
struct FaceSt {
    size_t p, m, up1[2], dn1[2], up2[2], dn2[2];
    double a, s1, s2, div;
};

__device__ FaceSt face_st(const Geo& g, const double* U, const double* V, const double* W, int i,
                          int j, int k, int fam)
{
    FaceSt f;
    const size_t o = g.at(i, j, k);
    if (fam == 0) {
        double vr = V[g.at(i + 1, j - 1, k)], vru = V[g.at(i + 1, j, k)], vu = V[o];
        double vc = V[g.at(i, j - 1, k)];
        double wr = W[g.at(i + 1, j, k - 1)], wrf = W[g.at(i + 1, j, k)], wf = W[o];
        double wc = W[g.at(i, j, k - 1)];
        f.a = U[o];
        f.s1 = 0.25 * (vr + vru + vu + vc);
        f.s2 = 0.25 * (wr + wrf + wf + wc);
        f.div = 0.25 * f.a *
                (U[g.at(i + 1, j, k)] - U[g.at(i - 1, j, k)] + vr - vru - vu + vc + wr - wrf - wf + wc);
        f.p = g.at(i + 1, j, k);
        f.m = o;
        f.up1[0] = g.at(i + 1, j + 1, k); f.up1[1] = g.at(i, j + 1, k);
        f.dn1[0] = g.at(i, j - 1, k); f.dn1[1] = g.at(i + 1, j - 1, k);
        f.up2[0] = g.at(i + 1, j, k + 1); f.up2[1] = g.at(i, j, k + 1);
        f.dn2[0] = g.at(i, j, k - 1); f.dn2[1] = g.at(i + 1, j, k - 1);
    } else if (fam == 1) {
        double uu = U[g.at(i - 1, j + 1, k)], uru = U[g.at(i, j + 1, k)], ur = U[o];
        double uc = U[g.at(i - 1, j, k)], ulu = U[g.at(i - 2, j + 1, k)], ul = U[g.at(i - 2, j, k)];
        double wu = W[g.at(i, j + 1, k - 1)], wuf = W[g.at(i, j + 1, k)], wf = W[o];
        double wc = W[g.at(i, j, k - 1)];
        f.a = V[o];
        f.s1 = 0.25 * (uu + uru + ur + uc);
        f.s2 = 0.25 * (wu + wuf + wf + wc);
        f.div = 0.25 * f.a *
                (V[g.at(i, j + 1, k)] - V[g.at(i, j - 1, k)] + wu - wuf - wf + wc + uu - ulu - ul + uc);
        f.p = g.at(i, j + 1, k);
        f.m = o;
        f.up1[0] = g.at(i + 1, j + 1, k); f.up1[1] = g.at(i + 1, j, k);
        f.dn1[0] = g.at(i - 1, j, k); f.dn1[1] = g.at(i - 1, j + 1, k);
        f.up2[0] = g.at(i, j + 1, k + 1); f.up2[1] = g.at(i, j, k + 1);
        f.dn2[0] = g.at(i, j, k - 1); f.dn2[1] = g.at(i, j + 1, k - 1);
    } else {
        double ur = U[o], urf = U[g.at(i, j, k + 1)], uf = U[g.at(i - 1, j, k + 1)];
        double uc = U[g.at(i - 1, j, k)];
        double vf = V[g.at(i, j - 1, k + 1)], vuf = V[g.at(i, j, k + 1)], vu = V[o];
        double vc = V[g.at(i, j - 1, k)];
        f.a = W[o];
        f.s1 = 0.25 * (ur + urf + uf + uc);
        f.s2 = 0.25 * (vf + vuf + vu + vc);
        f.div = 0.25 * f.a *
                (W[g.at(i, j, k + 1)] - W[g.at(i, j, k - 1)] + uf - urf - ur + uc + vf - vuf - vu + vc);
        f.p = g.at(i, j, k + 1);
        f.m = o;
        f.up1[0] = g.at(i + 1, j, k + 1); f.up1[1] = g.at(i + 1, j, k);
        f.dn1[0] = g.at(i - 1, j, k); f.dn1[1] = g.at(i - 1, j, k + 1);
        f.up2[0] = g.at(i, j + 1, k + 1); f.up2[1] = g.at(i, j + 1, k);
        f.dn2[0] = g.at(i, j - 1, k); f.dn2[1] = g.at(i, j - 1, k + 1);
    }
    return f;
}

// limited velocity of a face with unlimited velocity a between lower cell l, upper r
__device__ __forceinline__ double limited(double a, double bu_l, double bd_l, double bu_r,
                                          double bd_r)
{
    if (a > 0) {
        int w = pick(bd_l, bu_r);
        return (w == 0 ? 1.0 : (w == 1 ? bd_l : bu_r)) * a;
    }
    if (a < 0) {
        int w = pick(bu_l, bd_r);
        return (w == 0 ? 1.0 : (w == 1 ? bu_l : bd_r)) * a;
    }
    return 0.0;
}

// ------------------------------------------------------------------ adjoint parts ---

struct AdjBufs {
    const double *U, *V, *W, *c0, *c1, *A0, *A1, *A2, *bup, *bdn;
    const int* code;
    int nonosc;
};

__device__ __forceinline__ void nbrs(const Geo& g, size_t t, size_t lo[3], size_t up[3])
{
    int i, j, k;
    g.ijk(t, i, j, k);
    lo[0] = g.at(i - 1, j, k); up[0] = g.at(i + 1, j, k);
    lo[1] = g.at(i, j - 1, k); up[1] = g.at(i, j + 1, k);
    lo[2] = g.at(i, j, k - 1); up[2] = g.at(i, j, k + 1);
}

__device__ __forceinline__ double fam_a(const AdjBufs& b, int d, size_t o)
{
    return d == 0 ? b.A0[o] : (d == 1 ? b.A1[o] : b.A2[o]);
}

// Corrective donor pass and limiter part 1, transposed, in one pass over the cells:
//   l1 = lout + the donor gather over the six faces (limited velocities);
//   lam_Al of each face of the cell = gv * (lout[upper] - lout[lower]) (recomputed for
//   the lower faces rather than read back); lA = m * lam_Al on the upper faces, m the
//   limiter's multiplier; the multipliers' ratio adjoints, gathered from the six faces,
//   turned into bound and flux-sum adjoints (iu, id, lain, laout).  Without the limiter
//   lA = lam_Al and the ratio outputs are untouched.  lq += lout.
__global__ void k_adj_donor_limit(Geo g, AdjBufs b, const double* __restrict__ lout,
                                  double* __restrict__ l1, double* __restrict__ lA0,
                                  double* __restrict__ lA1, double* __restrict__ lA2,
                                  double* __restrict__ iu, double* __restrict__ id,
                                  double* __restrict__ lain, double* __restrict__ laout,
                                  double* __restrict__ lq)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= g.n) return;
    size_t lo[3], up[3];
    nbrs(g, t, lo, up);
    double* lA[3] = {lA0, lA1, lA2};
    const double lt = lout[t], ct = b.c1[t];
    double bu_t = 0.0, bd_t = 0.0;
    if (b.nonosc) {
        bu_t = b.bup[t];
        bd_t = b.bdn[t];
    }
    double acc = lt, su = 0.0, sd = 0.0;
    double fl[3], fu[3];  // the unlimited donor fluxes of the lower and upper faces
    for (int d = 0; d < 3; d++) {
        double gl, gr, gv;
        {  // upper face (t, r): t is the lower cell
            const size_t r = up[d];
            const double a = fam_a(b, d, t);
            fu[d] = 0.5 * F2(ct, b.c1[r], a);
            double au = a, m = 1.0;
            if (b.nonosc) {
                const double bu_r = b.bup[r], bd_r = b.bdn[r];
                if (a > 0) {
                    const int w = pick(bd_t, bu_r);
                    m = w == 0 ? 1.0 : (w == 1 ? bd_t : bu_r);
                } else if (a < 0) {
                    const int w = pick(bu_t, bd_r);
                    m = w == 0 ? 1.0 : (w == 1 ? bu_t : bd_r);
                } else {
                    m = 0.0;
                }
                au = limited(a, bu_t, bd_t, bu_r, bd_r);
            }
            fpart(ct, b.c1[r], au, gl, gr, gv);
            const double lf = lout[r] - lt;
            acc += gl * lf;
            const double L = gv * lf;
            if (b.nonosc) {
                if (a > 0 && pick(bd_t, b.bup[r]) == 1) sd += a * L;
                if (a < 0 && pick(bu_t, b.bdn[r]) == 1) su += a * L;
            }
            lA[d][t] = m * L;
        }
        {  // lower face (l, t): t is the upper cell
            const size_t l = lo[d];
            const double a = fam_a(b, d, l);
            fl[d] = 0.5 * F2(b.c1[l], ct, a);
            double al = a;
            double bu_l = 0.0, bd_l = 0.0;
            if (b.nonosc) {
                bu_l = b.bup[l];
                bd_l = b.bdn[l];
                al = limited(a, bu_l, bd_l, bu_t, bd_t);
            }
            fpart(b.c1[l], ct, al, gl, gr, gv);
            const double lf = lt - lout[l];
            acc += gr * lf;
            if (b.nonosc) {
                const double L = gv * lf;
                if (a > 0 && pick(bd_l, bu_t) == 2) su += a * L;
                if (a < 0 && pick(bu_l, bd_t) == 2) sd += a * L;
            }
        }
    }
    l1[t] = acc;
    if (lq) lq[t] += lt;
    if (!b.nonosc) return;
    // ratio adjoints -> bound / flux-sum adjoints (pointwise).  A ratio that was not
    // selected carries no adjoint and may be inf: skipped.  The flux sums are those the
    // forward formed from the same six fluxes.
    double ain, aout;
    flux_sums(fl[0], fu[0], fl[1], fu[1], fl[2], fu[2], ain, aout);
    double u = 0.0, dn = 0.0, li = 0.0, lo2 = 0.0;
    if (su != 0.0) {
        u = su / (ain + MINV);
        li = -u * bu_t;
    }
    if (sd != 0.0) {
        dn = sd / (aout + MINV);
        lo2 = -dn * bd_t;
    }
    iu[t] = u;
    id[t] = dn;
    lain[t] = li;
    laout[t] = lo2;
}

// limiter transposed, part 3 (gather): bounds -> c_in / c1; flux sums -> c1 and the
// unlimited faces.  lin receives the c_in part, l1 and lA accumulate.
__global__ void k_adj_limit3(Geo g, AdjBufs b, const double* __restrict__ iu,
                             const double* __restrict__ id, const double* __restrict__ lain,
                             const double* __restrict__ laout, double* __restrict__ lin,
                             double* __restrict__ l1, double* __restrict__ lA0,
                             double* __restrict__ lA1, double* __restrict__ lA2)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= g.n) return;
    size_t lo[3], up[3];
    nbrs(g, t, lo, up);
    double* lA[3] = {lA0, lA1, lA2};
    // neighbour y and the code q under which y sees t: y = t + e  =>  t is y's "-e".
    // A bound that is the bounded cell's own c1 value enters bup = (hi - c)/(..) twice
    // with opposite signs; that pair is folded (skipped) by cell identity, as in the
    // oracle, because each half alone can be ~1/DBL_MIN.
    const size_t ys[7] = {t, up[0], lo[0], up[1], lo[1], up[2], lo[2]};
    const int qs[7] = {0, 1, 2, 3, 4, 5, 6};
    double s0 = 0.0, s1 = 0.0;
    for (int m = 0; m < 7; m++) {
        size_t y = ys[m];
        int c = b.code[y], hc = c & 0xff, lc = (c >> 8) & 0xff;
        bool folds = (y == t);  // t is y itself: its own c1 value is folded
        if (hc >> 1 == qs[m] && !(folds && (hc & 1))) {
            if (hc & 1) s1 += iu[y]; else s0 += iu[y];
        }
        if (lc >> 1 == qs[m] && !(folds && (lc & 1))) {
            if (lc & 1) s1 -= id[y]; else s0 -= id[y];
        }
    }
    {
        const size_t nb[7] = {t, lo[0], up[0], lo[1], up[1], lo[2], up[2]};
        int c = b.code[t], hc = c & 0xff, lc = (c >> 8) & 0xff;
        bool hself = nb[hc >> 1] == t && (hc & 1), lself = nb[lc >> 1] == t && (lc & 1);
        if (!hself) s1 -= iu[t];
        if (!lself) s1 += id[t];
    }
    // flux sums: face (L, R) gets lam_f = (laout_L [f>0] - lain_L [f<0])
    //                                    + (lain_R [f>0] - laout_R [f<0])
    double acc = 0.0;
    for (int d = 0; d < 3; d++) {
        double gl, gr, gv;
        size_t r = up[d], l = lo[d];
        double a = fam_a(b, d, t);
        double f = 0.5 * F2(b.c1[t], b.c1[r], a);
        double lf = (f > 0 ? laout[t] : 0.0) - (f < 0 ? lain[t] : 0.0) +
                    (f > 0 ? lain[r] : 0.0) - (f < 0 ? laout[r] : 0.0);
        fpart(b.c1[t], b.c1[r], a, gl, gr, gv);
        acc += gl * lf;
        lA[d][t] += gv * lf;
        a = fam_a(b, d, l);
        f = 0.5 * F2(b.c1[l], b.c1[t], a);
        lf = (f > 0 ? laout[l] : 0.0) - (f < 0 ? lain[l] : 0.0) + (f > 0 ? lain[t] : 0.0) -
             (f < 0 ? laout[t] : 0.0);
        fpart(b.c1[l], b.c1[t], a, gl, gr, gv);
        acc += gr * lf;
    }
    lin[t] = s0;
    l1[t] += s1 + acc;
}

// Scatter variant of the transposed antidiffusive stage: each face adds its ten
// contributions to l1 with hardware FP64 atomics (summation order, and so the last bits,
// vary between runs).
__global__ void k_adj_anti_atomic(Geo g, AdjBufs b, const double* __restrict__ lA0,
                                  const double* __restrict__ lA1, const double* __restrict__ lA2,
                                  double* __restrict__ l1)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= g.n) return;
    int i, j, k;
    g.ijk(t, i, j, k);
    const double* lA[3] = {lA0, lA1, lA2};
    const double* c = b.c1;
    for (int fam = 0; fam < 3; fam++) {
        double L = lA[fam][t];
        if (L == 0.0) continue;
        FaceSt f = face_st(g, b.U, b.V, b.W, i, j, k, fam);
        double p = c[f.p], m = c[f.m];
        double K = ab(f.a) * (1.0 - ab(f.a));
        double den = p + m + MINV;
        double n1 = c[f.up1[0]] + c[f.up1[1]] - c[f.dn1[0]] - c[f.dn1[1]];
        double D1 = c[f.up1[0]] + c[f.up1[1]] + c[f.dn1[0]] + c[f.dn1[1]] + MINV;
        double n2 = c[f.up2[0]] + c[f.up2[1]] - c[f.dn2[0]] - c[f.dn2[1]];
        double D2 = c[f.up2[0]] + c[f.up2[1]] + c[f.dn2[0]] + c[f.dn2[1]] + MINV;
        double gb = -0.5 * f.a;
        double kp = L * (K * (((2.0 * m + MINV) / den) / den));
        double km = L * (-K * (((2.0 * p + MINV) / den) / den));
        double ku1 = L * (gb * f.s1 * (((D1 - n1) / D1) / D1));
        double kd1 = L * (gb * f.s1 * (((-D1 - n1) / D1) / D1));
        double ku2 = L * (gb * f.s2 * (((D2 - n2) / D2) / D2));
        double kd2 = L * (gb * f.s2 * (((-D2 - n2) / D2) / D2));
        unsafeAtomicAdd(l1 + f.p, kp);
        unsafeAtomicAdd(l1 + f.m, km);
        unsafeAtomicAdd(l1 + f.up1[0], ku1);
        unsafeAtomicAdd(l1 + f.up1[1], ku1);
        unsafeAtomicAdd(l1 + f.dn1[0], kd1);
        unsafeAtomicAdd(l1 + f.dn1[1], kd1);
        unsafeAtomicAdd(l1 + f.up2[0], ku2);
        unsafeAtomicAdd(l1 + f.up2[1], ku2);
        unsafeAtomicAdd(l1 + f.dn2[0], kd2);
        unsafeAtomicAdd(l1 + f.dn2[1], kd2);
    }
}

// ------------------------------ corrective donor pass + limiter transposed, tiled ---
//
// k_adj_donor_limit followed by k_adj_limit3 in one LDS-tiled pass (with the limiter).
// Iteration p: publish the plane-p inputs -> barrier -> stage 1 on the tile and ring 1 of
// plane p (the donor gather, lam_Al of the six faces, the ratio adjoints iu, id, lain,
// laout) -> barrier -> stage 2 on the tile of plane k = p - 1 (the bound gather over the
// 7-point neighbourhood and the flux-sum adjoints of the six faces).  LDS slots: lout,
// bup, bdn one plane (read in stage 1 only), c1, A0, A1 three (stage 1 of plane p, stage
// 2 of plane p - 1), the stage-1 results and the codes two (both written after the
// first barrier: stage 2 of the previous iteration reads the other slot).  A z face lies
// between two planes of one column, so both stages form it once: as the upper face of a
// plane, carrying the lower face's terms of the next plane in registers (zface1, zface2).
struct LimArgs {
    int nx, ny, nz;
    long long np;
    int tiles_x, tiles_y, zc;
    const double *lout, *c1, *A0, *A1, *A2, *bup, *bdn;
    const int* code;
    double *l1, *lA0, *lA1, *lA2, *lin, *lq;
};

// own-column values of one region cell: planes p-1, p, p+1 of lout, c1, the ratios; the
// z-face velocity below (p-1) and above (p); plane p of the rest; n*: the next plane
struct LimCol {
    double lo[3], c1[3], bu[3], bd[3], a2m, a2c, a0, a1;
    int cd;
    double nlo, nc1, nbu, nbd, na0, na1, na2;
    int ncd;
};

__device__ __forceinline__ int ld_code(const int* plane, unsigned colb)
{
    return *(const int*)((const char*)plane + colb / 2);
}

// The LDS planes of k_adj_lim_t.
template <int R>
struct LimLds {
    double lo[R], bu[R], bd[R], c[3][R], a0[3][R], a1[3][R];
    double iu[2][R], id[2][R], li[2][R], lx[2][R];
    int cd[2][R];
};

// Ring-2 cell j of a TX x TY tile's region without its four corners, which no stage
// reads (stage 1 on ring 1 reads the four in-plane neighbours only): the rows below and
// above ring 1, then the two side columns.
template <int TX, int TY>
__device__ __forceinline__ void ring2_nc(int j, int& rx, int& ry)
{
    if (j < 2 * (TX + 2)) {
        ry = j < TX + 2 ? 0 : TY + 3;
        rx = 1 + j % (TX + 2);
    } else {
        const int r = j - 2 * (TX + 2);
        ry = 1 + r / 2;
        rx = (r & 1) ? TX + 3 : 0;
    }
}

// One role's share of a work item.  TILE: lane = tile cell (stages 1 and 2, the stores).
// !TILE: lane of the ring wave, with its ring-1 cell (stage 1; the last lanes have none
// and load a duplicate column) and its ring-2 cell (loads and publishes only).  The two
// roles are separate code paths, so no thread holds the rolling registers of two
// columns and the tile path's stage 2 does not share its registers with ring-2 loads.
template <int TX, int TY, bool TILE>
__device__ __forceinline__ void lim_part(const LimArgs& a, int item, int lane,
                                         LimLds<tile::GeoB<TX, TY>::R>& sm)
{
    using G = tile::GeoB<TX, TY>;
    using tile::ld;
    using tile::st_nt;  // streaming: the adjoint's arrays far exceed the caches at its sizes
    constexpr int RX = G::RX;
    const int bx = item % a.tiles_x, by = (item / a.tiles_x) % a.tiles_y;
    const int bz = item / (a.tiles_x * a.tiles_y);
    // own cell (tile: stages 1 and 2; ring 1: stage 1) and ring-2 cell (loads)
    int ox, oy, sx = 0, sy = 0;
    bool comp = true;
    if constexpr (TILE) {
        ox = 2 + lane % TX;
        oy = 2 + lane / TX;
    } else {
        comp = lane < G::H1;
        G::G::halo(comp ? lane : 0, ox, oy);
        ring2_nc<TX, TY>(lane, sx, sy);
    }
    constexpr bool tcell = TILE, sec = !TILE;
    const int L = oy * RX + ox, S = sy * RX + sx;
    unsigned colb = tile::column<TX, TY>(a.nx, a.ny, bx, by, ox, oy);
    unsigned scolb = tile::column<TX, TY>(a.nx, a.ny, bx, by, sx, sy);
    const bool store = tcell && bx * TX + ox - 2 < a.nx && by * TY + oy - 2 < a.ny;
    const int z0 = bz * a.zc;
    const int z1 = min(z0 + a.zc, a.nz);
    // planes p + 2, p + 1, p, p - 1, p - 2 of iteration p
    tile::ZCursor zp2(z0 + 1, a.nz, a.np), zp1(z0, a.nz, a.np), zp0(z0 - 1, a.nz, a.np);
    tile::ZCursor zm1(z0 - 2, a.nz, a.np);

    // loads of a column: everything for a tile or ring-1 cell, the in-plane values for a
    // ring-2 cell
    auto load_next = [&](LimCol& q, unsigned cb, bool full, const tile::ZCursor& zn,
                         const tile::ZCursor& zc1) {
        q.nlo = ld(a.lout + zn.off, cb);
        q.nc1 = ld(a.c1 + zn.off, cb);
        q.nbu = ld(a.bup + zn.off, cb);
        q.nbd = ld(a.bdn + zn.off, cb);
        q.na0 = ld(a.A0 + zc1.off, cb);
        q.na1 = ld(a.A1 + zc1.off, cb);
        if (full) {
            q.na2 = ld(a.A2 + zc1.off, cb);
            q.ncd = ld_code(a.code + zc1.off, cb);
        }
    };
    auto init = [&](LimCol& q, unsigned cb, bool full) {
        q.lo[0] = ld(a.lout + zm1.off, cb);
        q.lo[1] = ld(a.lout + zp0.off, cb);
        q.c1[0] = ld(a.c1 + zm1.off, cb);
        q.c1[1] = ld(a.c1 + zp0.off, cb);
        q.bu[0] = ld(a.bup + zm1.off, cb);
        q.bu[1] = ld(a.bup + zp0.off, cb);
        q.bd[0] = ld(a.bdn + zm1.off, cb);
        q.bd[1] = ld(a.bdn + zp0.off, cb);
        q.a0 = ld(a.A0 + zp0.off, cb);
        q.a1 = ld(a.A1 + zp0.off, cb);
        q.a2m = q.a2c = 0.0;
        q.cd = 0;
        if (full) {
            q.a2m = ld(a.A2 + zm1.off, cb);
            q.a2c = ld(a.A2 + zp0.off, cb);
            q.cd = ld_code(a.code + zp0.off, cb);
        }
        // plane p + 1 of the first iteration, then the prefetch of the second
        q.lo[2] = ld(a.lout + zp1.off, cb);
        q.c1[2] = ld(a.c1 + zp1.off, cb);
        q.bu[2] = ld(a.bup + zp1.off, cb);
        q.bd[2] = ld(a.bdn + zp1.off, cb);
        load_next(q, cb, full, zp2, zp1);
    };
    auto rotate = [&](LimCol& q) {
        q.lo[0] = q.lo[1];
        q.lo[1] = q.lo[2];
        q.lo[2] = q.nlo;
        q.c1[0] = q.c1[1];
        q.c1[1] = q.c1[2];
        q.c1[2] = q.nc1;
        q.bu[0] = q.bu[1];
        q.bu[1] = q.bu[2];
        q.bu[2] = q.nbu;
        q.bd[0] = q.bd[1];
        q.bd[1] = q.bd[2];
        q.bd[2] = q.nbd;
        q.a2m = q.a2c;
        q.a2c = q.na2;
        q.a0 = q.na0;
        q.a1 = q.na1;
        q.cd = q.ncd;
    };
    LimCol P, H;
    init(P, colb, true);
    if (sec) init(H, scolb, false);

    // tile cell: stage-1 results of planes p-2, p-1 (index 0, 1; 2 = plane p), the codes,
    // the carried donor sum and lam_Al * m of plane p-1
    double siu0 = 0, sid0 = 0, dacc_m = 0.0, lam_m[3] = {0, 0, 0};
    // stage 2: the term gr*lf of the z face below the cell of plane k (its upper z face
    // at plane k - 1)
    double z2 = 0.0;
    int scd0 = 0;
    // the emission adjoint of plane p, loaded one iteration ahead
    double lq_p = 0.0;
    // stage 1: the z face below the own cell of plane p, formed as the upper z face of
    // plane p - 1 (first: the face between planes z0 - 2 and z0 - 1, from the loads)
    ZFace zc = zface1(P.c1[0], P.c1[1], P.a2m, P.lo[0], P.lo[1], P.bu[0], P.bd[0], P.bu[1],
                      P.bd[1]);

    for (int pl = z0 - 1; pl <= z1; pl++) {
        const int p = __builtin_amdgcn_readfirstlane(pl);
        tile::keep32(colb);  // scalar-base addressing (see keep32)
        tile::keep32(scolb);
        const int s3 = (p + 3) % 3, k3 = (p + 2) % 3;  // slots of planes p, p-1 (c1, A0, A1)
        const int s2 = (p + 2) % 2, k2 = (p + 1) % 2;  // slots of planes p, p-1 (stage 1)
        if (comp) {
            sm.lo[L] = P.lo[1];
            sm.bu[L] = P.bu[1];
            sm.bd[L] = P.bd[1];
            sm.c[s3][L] = P.c1[1];
            sm.a0[s3][L] = P.a0;
            sm.a1[s3][L] = P.a1;
        }
        if (sec) {
            sm.lo[S] = H.lo[1];
            sm.bu[S] = H.bu[1];
            sm.bd[S] = H.bd[1];
            sm.c[s3][S] = H.c1[1];
            sm.a0[s3][S] = H.a0;
            sm.a1[s3][S] = H.a1;
        }
        tile::lds_barrier();
        // stage 1 on cell M of plane p (the merged k_adj_donor_limit, same operation order)
        auto stage1 = [&](int M, const LimCol& q, double& dacc, double lam[3], double& iu,
                          double& id, double& li, double& lx, const ZFace& zc, ZFace& zn) {
            const double* C = sm.c[s3];
            const double lt = q.lo[1], ct = q.c1[1], bu_t = q.bu[1], bd_t = q.bd[1];
            double acc = lt, su = 0.0, sd = 0.0;
            double fl[3], fu[3];  // the unlimited donor fluxes of the lower / upper faces
#pragma unroll
            for (int d = 0; d < 2; d++) {
                double gl, gr, gv;
                {  // upper face (M, r)
                    double a_, lo_r, c_r, bu_r, bd_r;
                    if (d == 0) {
                        a_ = q.a0; lo_r = sm.lo[M + 1]; c_r = C[M + 1]; bu_r = sm.bu[M + 1]; bd_r = sm.bd[M + 1];
                    } else {
                        a_ = q.a1; lo_r = sm.lo[M + RX]; c_r = C[M + RX]; bu_r = sm.bu[M + RX]; bd_r = sm.bd[M + RX];
                    }
                    // limiter multiplier without branches: the ratio pair of the
                    // flow direction, the selection of mn2(1, mn2(x, y))
                    fu[d] = 0.5 * F2(ct, c_r, a_);
                    const bool pf = a_ > 0, nf = a_ < 0;
                    const double x = pf ? bd_t : bu_t, y = pf ? bu_r : bd_r;
                    const int w = pick(x, y);
                    const double m = (pf || nf) ? (w == 0 ? 1.0 : (w == 1 ? x : y)) : 0.0;
                    const double au = (pf || nf) ? m * a_ : 0.0;
                    fpart(ct, c_r, au, gl, gr, gv);
                    const double lf = lo_r - lt;
                    acc += gl * lf;
                    const double Lf = gv * lf;
                    const double sel = (w == 1 && (pf || nf)) ? a_ * Lf : 0.0;
                    sd += pf ? sel : 0.0;
                    su += nf ? sel : 0.0;
                    lam[d] = m * Lf;
                }
                {  // lower face (l, M)
                    double a_, lo_l, c_l, bu_l, bd_l;
                    if (d == 0) {
                        a_ = sm.a0[s3][M - 1]; lo_l = sm.lo[M - 1]; c_l = C[M - 1]; bu_l = sm.bu[M - 1]; bd_l = sm.bd[M - 1];
                    } else {
                        a_ = sm.a1[s3][M - RX]; lo_l = sm.lo[M - RX]; c_l = C[M - RX]; bu_l = sm.bu[M - RX]; bd_l = sm.bd[M - RX];
                    }
                    fl[d] = 0.5 * F2(c_l, ct, a_);
                    const bool pf = a_ > 0, nf = a_ < 0;
                    const double x = pf ? bd_l : bu_l, y = pf ? bu_t : bd_t;
                    const int w = pick(x, y);
                    const double al =
                        (pf || nf) ? (w == 0 ? 1.0 : (w == 1 ? x : y)) * a_ : 0.0;
                    fpart(c_l, ct, al, gl, gr, gv);
                    const double lf = lt - lo_l;
                    acc += gr * lf;
                    const double Lf = gv * lf;
                    const double sel = (w == 2 && (pf || nf)) ? a_ * Lf : 0.0;
                    su += pf ? sel : 0.0;
                    sd += nf ? sel : 0.0;
                }
            }
            {  // the upper z face, both sides at once (the upper cell's share is carried
               // to the next plane), then the lower z face: the upper z face of the
               // plane below
                zn = zface1(ct, q.c1[2], q.a2c, lt, q.lo[2], bu_t, bd_t, q.bu[2], q.bd[2]);
                fu[2] = zn.f;
                acc += zn.glf;
                sd += zn.sdu;
                su += zn.suu;
                lam[2] = zn.lam;
                fl[2] = zc.f;
                acc += zc.grlf;
                su += zc.sul;
                sd += zc.sdl;
            }
            dacc = acc;
            // the flux sums of the forward, formed again from the same six fluxes
            double ain, aout;
            flux_sums(fl[0], fu[0], fl[1], fu[1], fl[2], fu[2], ain, aout);
            iu = id = li = lx = 0.0;
            if (su != 0.0) {
                iu = su / (ain + MINV);
                li = -iu * bu_t;
            }
            if (sd != 0.0) {
                id = sd / (aout + MINV);
                lx = -id * bd_t;
            }
        };
        double dacc = 0.0, lam[3] = {0.0, 0.0, 0.0};
        if (comp) {
            double iu, id, li, lx;
            ZFace zn;
            stage1(L, P, dacc, lam, iu, id, li, lx, zc, zn);
            zc = zn;
            // the codes of plane p go to LDS here, after the first barrier, like the
            // stage-1 results: their slot held plane p - 2, which stage 2 of the previous
            // iteration may still be reading until every wave has reached that barrier
            sm.cd[s2][L] = P.cd;
            sm.iu[s2][L] = iu;
            sm.id[s2][L] = id;
            sm.li[s2][L] = li;
            sm.lx[s2][L] = lx;
        }
        if (store && p >= z0 && p < z1 && a.lq) st_nt(a.lq + zp0.off, colb, lq_p + P.lo[1]);
        tile::lds_barrier();
        // stage 2 on the tile cell of plane k = p - 1 (k_adj_limit3, same operation order)
        const int k = p - 1;
        if (tcell && k >= z0 && k < z1) {
            const double* C = sm.c[k3];
            const double *IU = sm.iu[k2], *ID = sm.id[k2], *LI = sm.li[k2], *LX = sm.lx[k2];
            const int* CD = sm.cd[k2];
            const double ct = C[L];  // c1 of plane k
            // the cell's own column: stage-1 results and codes of planes k (LDS slot k2),
            // k + 1 (slot s2) and k - 1 (registers)
            const double iu_k = IU[L], id_k = ID[L], iu_n = sm.iu[s2][L], id_n = sm.id[s2][L];
            const int cd_k = CD[L], cd_n = sm.cd[s2][L];
            // neighbour y, the code q under which y sees t, whether y is t itself (extents
            // of 1 wrap a neighbour onto the cell)
            double s0 = 0.0, s1 = 0.0;
            const bool fx = a.nx == 1, fy = a.ny == 1, fz = a.nz == 1;
            // (selects rather than branches: a skipped term adds zero, which changes
            // nothing but the sign of a zero sum)
            auto bound = [&](int cq, double iuy, double idy, int q, bool folds) {
                const int hc = cq & 0xff, lc = (cq >> 8) & 0xff;
                const bool hm = hc >> 1 == q && !(folds && (hc & 1));
                const bool lm = lc >> 1 == q && !(folds && (lc & 1));
                s0 += (hm && !(hc & 1)) ? iuy : 0.0;
                s1 += (hm && (hc & 1)) ? iuy : 0.0;
                s0 -= (lm && !(lc & 1)) ? idy : 0.0;
                s1 -= (lm && (lc & 1)) ? idy : 0.0;
            };
            bound(cd_k, iu_k, id_k, 0, true);
            bound(CD[L + 1], IU[L + 1], ID[L + 1], 1, fx);
            bound(CD[L - 1], IU[L - 1], ID[L - 1], 2, fx);
            bound(CD[L + RX], IU[L + RX], ID[L + RX], 3, fy);
            bound(CD[L - RX], IU[L - RX], ID[L - RX], 4, fy);
            bound(cd_n, iu_n, id_n, 5, fz);
            bound(scd0, siu0, sid0, 6, fz);
            {
                const int c = cd_k, hc = c & 0xff, lc = (c >> 8) & 0xff;
                auto self = [&](int q) {
                    return q == 0 || (q <= 2 && fx) || ((q == 3 || q == 4) && fy) || (q >= 5 && fz);
                };
                const bool hself = self(hc >> 1) && (hc & 1), lself = self(lc >> 1) && (lc & 1);
                if (!hself) s1 -= iu_k;
                if (!lself) s1 += id_k;
            }
            // flux sums of the six faces
            const double li_t = LI[L], lx_t = LX[L];
            double acc = 0.0, lAf[3], z2n;
#pragma unroll
            for (int d = 0; d < 2; d++) {
                double gl, gr, gv;
                {  // upper face (t, r)
                    double a_, c_r, li_r, lx_r;
                    if (d == 0) {
                        a_ = sm.a0[k3][L]; c_r = C[L + 1]; li_r = LI[L + 1]; lx_r = LX[L + 1];
                    } else {
                        a_ = sm.a1[k3][L]; c_r = C[L + RX]; li_r = LI[L + RX]; lx_r = LX[L + RX];
                    }
                    // (lx_t [f>0] - li_t [f<0]) + (li_r [f>0] - lx_r [f<0]), the zero
                    // terms dropped (the same value up to the sign of a zero)
                    const double f = 0.5 * F2(ct, c_r, a_);
                    const double lf = f > 0 ? lx_t + li_r : (f < 0 ? -(li_t + lx_r) : 0.0);
                    fpart(ct, c_r, a_, gl, gr, gv);
                    acc += gl * lf;
                    lAf[d] = lam_m[d] + gv * lf;
                }
                {  // lower face (l, t)
                    double a_, c_l, li_l, lx_l;
                    if (d == 0) {
                        a_ = sm.a0[k3][L - 1]; c_l = C[L - 1]; li_l = LI[L - 1]; lx_l = LX[L - 1];
                    } else {
                        a_ = sm.a1[k3][L - RX]; c_l = C[L - RX]; li_l = LI[L - RX]; lx_l = LX[L - RX];
                    }
                    const double f = 0.5 * F2(c_l, ct, a_);
                    const double lf = f > 0 ? lx_l + li_t : (f < 0 ? -(li_l + lx_t) : 0.0);
                    fpart(c_l, ct, a_, gl, gr, gv);
                    acc += gr * lf;
                }
            }
            {  // the upper z face (its gr*lf carried to the next plane), then the lower
               // z face: the upper z face of plane k - 1
                double glf, gvlf;
                zface2(ct, sm.c[s3][L], P.a2m, li_t, lx_t, sm.li[s2][L], sm.lx[s2][L], glf,
                       gvlf, z2n);
                acc += glf;
                lAf[2] = lam_m[2] + gvlf;
                acc += z2;
            }
            if (store) {
                st_nt(a.lin + zm1.off, colb, s0);
                st_nt(a.l1 + zm1.off, colb, dacc_m + (s1 + acc));
                st_nt(a.lA0 + zm1.off, colb, lAf[0]);
                st_nt(a.lA1 + zm1.off, colb, lAf[1]);
                st_nt(a.lA2 + zm1.off, colb, lAf[2]);
            }
            z2 = z2n;
        } else if (tcell && k == z0 - 1) {
            // the z face between planes z0 - 1 and z0: the lower face of the first plane
            double glf, gvlf;
            zface2(sm.c[k3][L], sm.c[s3][L], P.a2m, sm.li[k2][L], sm.lx[k2][L], sm.li[s2][L],
                   sm.lx[s2][L], glf, gvlf, z2);
        }
        // shift the tile cell's history by one plane: plane p - 1 becomes p - 2 (its LDS
        // slot is overwritten by the next iteration's stage 1)
        if constexpr (TILE) {
            siu0 = sm.iu[k2][L];
            sid0 = sm.id[k2][L];
            scd0 = sm.cd[k2][L];
            dacc_m = dacc;
            lam_m[0] = lam[0];
            lam_m[1] = lam[1];
            lam_m[2] = lam[2];
        }
        rotate(P);
        if (sec) rotate(H);
        zp2.next();
        zp1.next();
        zp0.next();
        zm1.next();
        if (p + 2 <= z1) {  // for the rotation after the next iteration: fields at p + 3,
                            // faces at p + 2 (the cursors now stand one plane higher)
            load_next(P, colb, true, zp2, zp1);
            if (sec) load_next(H, scolb, false, zp2, zp1);
        }
        if (store && p + 1 >= z0 && p + 1 < z1 && a.lq) lq_p = ld(a.lq + zp0.off, colb);
    }
}

// One workgroup per (tile, z-chunk): NT/64 tile waves and one ring wave.  The role
// numbering rotates by whole waves from item to item, so the ring wave (lighter: stage 1
// only) falls on a different SIMD in the workgroups that share a CU.  The LDS (53,760 B
// at 16x12) and the tile path's registers allow three workgroups per CU.
template <int TX, int TY>
__global__ void __launch_bounds__((tile::GeoB<TX, TY>::NTH)) __attribute__((amdgpu_waves_per_eu(3)))
k_adj_lim_t(LimArgs a)
{
    using G = tile::GeoB<TX, TY>;
    static_assert(G::NTH == G::NT + 64 && G::H1 <= 64, "one ring wave holds ring 1");
    static_assert(2 * (TX + 2) + 2 * (TY + 2) == 64, "one ring-2 cell per ring lane");
    __shared__ LimLds<G::R> sm;
    const int item = tile::xcd_item(blockIdx.x, gridDim.x);
    const int t = __builtin_amdgcn_readfirstlane(G::role(item)) + (int)(threadIdx.x & 63);
    if (t < G::NT)
        lim_part<TX, TY, true>(a, item, t, sm);
    else
        lim_part<TX, TY, false>(a, item, t - G::NT, sm);
}

int lim_tiled(const mpd_grid_s* g, const double* const in[6], const int* code,
              const double* lout, double* l1, double* const lA[3], double* lin, double* lq,
              hipStream_t s)
{
    constexpr int TX = 16, TY = 12;
    LimArgs a{};
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.np = (long long)g->nx * g->ny;
    a.tiles_x = (g->nx + TX - 1) / TX;
    a.tiles_y = (g->ny + TY - 1) / TY;
    const long long cols = (long long)a.tiles_x * a.tiles_y;
    int zc = g->nz;
    while (zc > 8 && cols * ((g->nz + zc - 1) / zc) < 2048) zc = (zc + 1) / 2;
    a.zc = zc;
    const int nchunks = (g->nz + zc - 1) / zc;
    a.lout = lout;
    a.c1 = in[0];
    a.A0 = in[1];
    a.A1 = in[2];
    a.A2 = in[3];
    a.bup = in[4];
    a.bdn = in[5];
    a.code = code;
    a.l1 = l1;
    a.lA0 = lA[0];
    a.lA1 = lA[1];
    a.lA2 = lA[2];
    a.lin = lin;
    a.lq = lq;
    hipLaunchKernelGGL((k_adj_lim_t<TX, TY>), dim3((unsigned)(cols * nchunks)),
                       dim3(tile::GeoB<TX, TY>::NTH), 0, s, a);
    MPD_HIP(hipGetLastError());
    return MPD_OK;
}

// ---------------------------------------- antidiffusive velocities transposed, tiled ---
//
// For every face the derivatives of its antidiffusive velocity (uu.c:18-32 pattern) with
// respect to its ten cell values, times the face adjoint L:
//   kp, km     the cells above / below the face:  +-K (2 m|p + MINV) / den^2,
//              K = |a| (1 - |a|), den = p + m + MINV
//   ku1, kd1   the two upper / two lower cells of the first transverse ratio n1/D1:
//              -a/2 s1 (+-D1 - n1) / D1^2;  ku2, kd2 likewise for the second ratio.
// Each cell then sums the coefficients of the faces whose stencils contain it (the
// stencils of face_st read backwards: 30 terms from the faces of planes k-1, k, k+1).
struct FaceK {
    double kp, km, ku1, kd1, ku2, kd2;
};

// 1/d for d >= DBL_MIN: hardware reciprocal estimate (relative error ~2^-24) and two
// Newton-Raphson steps r <- r + r(1 - d r): about one ulp.
__device__ __forceinline__ double recip(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ __forceinline__ FaceK face_k(double L, double a, double p, double m, double s1,
                                        double u1a, double u1b, double d1a, double d1b,
                                        double s2, double u2a, double u2b, double d2a,
                                        double d2b)
{
    // With the sums U = up pair, Dn = down pair of a ratio term, n = U - Dn and
    // D = U + Dn + MINV: D - n = 2 Dn + MINV and -D - n = -(2 U + MINV), formed directly
    // (no cancellation), and the factors L K and L (-a/2) s shared by each pair of
    // coefficients; every (x r) r keeps 1/den^2 from overflowing at den = MINV.
    FaceK k;
    const double LK = L * (ab(a) * (1.0 - ab(a)));
    const double r0 = recip(p + m + MINV);
    k.kp = LK * (((2.0 * m + MINV) * r0) * r0);
    k.km = -LK * (((2.0 * p + MINV) * r0) * r0);
    const double Lg = L * (-0.5 * a);
    const double U1 = u1a + u1b, Dn1 = d1a + d1b;
    const double r1 = recip((U1 + Dn1) + MINV);
    const double w1 = Lg * s1;
    k.ku1 = w1 * (((2.0 * Dn1 + MINV) * r1) * r1);
    k.kd1 = -w1 * (((2.0 * U1 + MINV) * r1) * r1);
    const double U2 = u2a + u2b, Dn2 = d2a + d2b;
    const double r2 = recip((U2 + Dn2) + MINV);
    const double w2 = Lg * s2;
    k.ku2 = w2 * (((2.0 * Dn2 + MINV) * r2) * r2);
    k.kd2 = -w2 * (((2.0 * U2 + MINV) * r2) * r2);
    return k;
}

struct AntiArgs {
    int nx, ny, nz;
    long long np;
    int tiles_x, tiles_y, zc;
    const double *c1, *U, *V, *W, *lA0, *lA1, *lA2;
    double* l1;  // the coefficient sums are added to it on every cell
};

// One workgroup per (tile, z-chunk), 2.5-D streaming as the forward kernels (mpd_tile.hpp):
// each thread owns a tile column and at most one halo column.  Iteration p: publish c1
// (p+1), U, V (p+1), W (p) -> barrier -> coefficients of the faces of plane p on the tile
// and ring 1; the 14 that other columns gather go to LDS -> barrier -> each tile cell
// adds the terms of plane-p faces to its running sums for planes p-1, p, p+1; the sum
// of plane p-1 is then complete.  LDS slots: c1 three planes, U, V, W two, the
// coefficients one (written after the first barrier of an iteration, read before the
// first barrier of the next).
template <int TX, int TY>
__global__ void __launch_bounds__((tile::GeoB<TX, TY>::NTH)) k_adj_anti_t(AntiArgs a)
{
    using G = tile::GeoB<TX, TY>;
    using tile::ld;
    using tile::st_nt;  // streaming: the adjoint's arrays far exceed the caches at its sizes
    constexpr int RX = G::RX, R = G::R;
    __shared__ double s_c[3][R], s_u[2][R], s_v[2][R], s_w[2][R];
    __shared__ double s_k[14][R];

    const int item = tile::xcd_item(blockIdx.x, gridDim.x);
    const int t = G::role(item);
    const int bx = item % a.tiles_x, by = (item / a.tiles_x) % a.tiles_y;
    const int bz = item / (a.tiles_x * a.tiles_y);
    // own cell (tile: coefficients + gather; ring 1: coefficients) and ring-2 cell (loads)
    int ox, oy, sx, sy;
    const int role = G::own(t, ox, oy);
    const bool comp = role >= 0, tcell = role == 0;
    const bool sec = G::second(t, sx, sy);
    const int L = oy * RX + ox, S = sy * RX + sx;
    unsigned colb = tile::column<TX, TY>(a.nx, a.ny, bx, by, ox, oy);
    unsigned scolb = tile::column<TX, TY>(a.nx, a.ny, bx, by, sx, sy);
    const bool store = tcell && bx * TX + ox - 2 < a.nx && by * TY + oy - 2 < a.ny;
    const int z0 = bz * a.zc;
    const int z1 = min(z0 + a.zc, a.nz);
    // planes p + 2, p + 1, p, p - 1 of iteration p
    tile::ZCursor zp2(z0 + 1, a.nz, a.np), zp1(z0, a.nz, a.np), zp0(z0 - 1, a.nz, a.np);
    tile::ZCursor zm1(z0 - 2, a.nz, a.np);

    // planes before the first iteration: c1 (z0-2, z0-1), U, V (z0-1), W (z0-2)
    {
        const int cs0 = (z0 + 1) % 3, cs1 = (z0 + 2) % 3, us = (z0 + 1) % 2, ws = z0 % 2;
        if (comp) {
            s_c[cs0][L] = ld(a.c1 + zm1.off, colb);
            s_c[cs1][L] = ld(a.c1 + zp0.off, colb);
            s_u[us][L] = ld(a.U + zp0.off, colb);
            s_v[us][L] = ld(a.V + zp0.off, colb);
            s_w[ws][L] = ld(a.W + zm1.off, colb);
        }
        if (sec) {
            s_c[cs0][S] = ld(a.c1 + zm1.off, scolb);
            s_c[cs1][S] = ld(a.c1 + zp0.off, scolb);
            s_u[us][S] = ld(a.U + zp0.off, scolb);
            s_v[us][S] = ld(a.V + zp0.off, scolb);
            s_w[ws][S] = ld(a.W + zm1.off, scolb);
        }
    }
    // values published / consumed by the next iteration: c1, U, V at p+1, W and the face
    // adjoints at p (own and ring-2 column), l1 at p (tile column)
    double nc = 0, nu = 0, nv = 0, nw = 0, nl0 = 0, nl1 = 0, nl2 = 0, nl1c = 0;
    double snc = 0, snu = 0, snv = 0, snw = 0;
    if (comp) {
        nc = ld(a.c1 + zp1.off, colb);
        nu = ld(a.U + zp1.off, colb);
        nv = ld(a.V + zp1.off, colb);
        nw = ld(a.W + zp0.off, colb);
        nl0 = ld(a.lA0 + zp0.off, colb);
        nl1 = ld(a.lA1 + zp0.off, colb);
        nl2 = ld(a.lA2 + zp0.off, colb);
        if (tcell) nl1c = ld(a.l1 + zp0.off, colb);
    }
    if (sec) {
        snc = ld(a.c1 + zp1.off, scolb);
        snu = ld(a.U + zp1.off, scolb);
        snv = ld(a.V + zp1.off, scolb);
        snw = ld(a.W + zp0.off, scolb);
    }
    double accm = 0.0, accc = 0.0, accn = 0.0, l1m = 0.0;
    for (int pl = z0 - 1; pl <= z1; pl++) {
        const int p = __builtin_amdgcn_readfirstlane(pl);
        tile::keep32(colb);  // scalar-base addressing (see keep32)
        tile::keep32(scolb);
        const int cm = (p + 2) % 3, cc = (p + 3) % 3, cq = (p + 4) % 3;  // c1 p-1, p, p+1
        const int up = (p + 2) % 2, uq = (p + 3) % 2;                     // U, V p, p+1
        const int wm = (p + 1) % 2, wc = (p + 2) % 2;                     // W p-1, p
        const double la0 = nl0, la1 = nl1, la2 = nl2, l1c = nl1c;
        if (comp) {
            s_c[cq][L] = nc;
            s_u[uq][L] = nu;
            s_v[uq][L] = nv;
            s_w[wc][L] = nw;
        }
        if (sec) {
            s_c[cq][S] = snc;
            s_u[uq][S] = snu;
            s_v[uq][S] = snv;
            s_w[wc][S] = snw;
        }
        if (p < z1) {
            if (comp) {
                nc = ld(a.c1 + zp2.off, colb);
                nu = ld(a.U + zp2.off, colb);
                nv = ld(a.V + zp2.off, colb);
                nw = ld(a.W + zp1.off, colb);
                nl0 = ld(a.lA0 + zp1.off, colb);
                nl1 = ld(a.lA1 + zp1.off, colb);
                nl2 = ld(a.lA2 + zp1.off, colb);
                if (tcell) nl1c = ld(a.l1 + zp1.off, colb);
            }
            if (sec) {
                snc = ld(a.c1 + zp2.off, scolb);
                snu = ld(a.U + zp2.off, scolb);
                snv = ld(a.V + zp2.off, scolb);
                snw = ld(a.W + zp1.off, scolb);
            }
        }
        tile::lds_barrier();
        // coefficients of the three faces above/right of the own cell in plane p
        FaceK kx{}, ky{}, kz{};
        if (comp) {
            const double *C = s_c[cc], *Cm = s_c[cm], *Cq = s_c[cq];
            const double *Up = s_u[up], *Uq = s_u[uq], *Vp = s_v[up], *Vq = s_v[uq];
            const double *Wm = s_w[wm], *Wc = s_w[wc];
            const int M = L;
            const double m = C[M];
            {  // x-face (uu.c:81-87)
                const double s1 = 0.25 * (Vp[M + 1 - RX] + Vp[M + 1] + Vp[M] + Vp[M - RX]);
                const double s2 = 0.25 * (Wm[M + 1] + Wc[M + 1] + Wc[M] + Wm[M]);
                kx = face_k(la0, Up[M], C[M + 1], m, s1, C[M + 1 + RX], C[M + RX], C[M - RX],
                            C[M + 1 - RX], s2, Cq[M + 1], Cq[M], Cm[M], Cm[M + 1]);
            }
            {  // y-face (vv.c:85-92)
                const double s1 = 0.25 * (Up[M - 1 + RX] + Up[M + RX] + Up[M] + Up[M - 1]);
                const double s2 = 0.25 * (Wm[M + RX] + Wc[M + RX] + Wc[M] + Wm[M]);
                ky = face_k(la1, Vp[M], C[M + RX], m, s1, C[M + 1 + RX], C[M + 1], C[M - 1],
                            C[M - 1 + RX], s2, Cq[M + RX], Cq[M], Cm[M], Cm[M + RX]);
            }
            {  // z-face (ww.c:248-255)
                const double s1 = 0.25 * (Up[M] + Uq[M] + Uq[M - 1] + Up[M - 1]);
                const double s2 = 0.25 * (Vq[M - RX] + Vq[M] + Vp[M] + Vp[M - RX]);
                kz = face_k(la2, Wc[M], Cq[M], m, s1, Cq[M + 1], C[M + 1], C[M - 1], Cq[M - 1],
                            s2, Cq[M + RX], C[M + RX], C[M - RX], Cq[M - RX]);
            }
            s_k[0][M] = kx.kp;
            s_k[1][M] = kx.ku1;
            s_k[2][M] = kx.kd1;
            s_k[3][M] = kx.ku2;
            s_k[4][M] = kx.kd2;
            s_k[5][M] = ky.kp;
            s_k[6][M] = ky.ku1;
            s_k[7][M] = ky.kd1;
            s_k[8][M] = ky.ku2;
            s_k[9][M] = ky.kd2;
            s_k[10][M] = kz.ku1;
            s_k[11][M] = kz.kd1;
            s_k[12][M] = kz.ku2;
            s_k[13][M] = kz.kd2;
        }
        tile::lds_barrier();
        if (tcell) {
            // terms of the faces of plane p, by target plane (own-column terms from
            // registers)
            accm += (kx.kd2 + s_k[4][L - 1]) + (ky.kd2 + s_k[9][L - RX]);
            accc += kx.km + s_k[0][L - 1] + s_k[1][L - 1 - RX] + s_k[1][L - RX] +
                    s_k[2][L + RX] + s_k[2][L - 1 + RX];
            accc += ky.km + s_k[5][L - RX] + s_k[6][L - 1 - RX] + s_k[6][L - 1] + s_k[7][L + 1] +
                    s_k[7][L + 1 - RX];
            const double zu1 = s_k[10][L - 1], zd1 = s_k[11][L + 1], zu2 = s_k[12][L - RX];
            const double zd2 = s_k[13][L + RX];
            accc += kz.km + zu1 + zd1 + zu2 + zd2;
            accn += (kx.ku2 + s_k[3][L - 1]) + (ky.ku2 + s_k[8][L - RX]) + kz.kp + zu1 + zd1 +
                    zu2 + zd2;
            if (store && p - 1 >= z0 && p - 1 < z1) st_nt(a.l1 + zm1.off, colb, l1m + accm);
            accm = accc;
            accc = accn;
            accn = 0.0;
            l1m = l1c;
        }
        zp2.next();
        zp1.next();
        zp0.next();
        zm1.next();
    }
}

int anti_tiled(const mpd_grid_s* g, const double* c1, double* const lA[3], double* l1,
               hipStream_t s)
{
    constexpr int TX = 16, TY = 10;
    AntiArgs a{};
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.np = (long long)g->nx * g->ny;
    a.tiles_x = (g->nx + TX - 1) / TX;
    a.tiles_y = (g->ny + TY - 1) / TY;
    // z-chunks: enough workgroups to fill the chip a few times over, at least 8 planes
    const long long cols = (long long)a.tiles_x * a.tiles_y;
    int zc = g->nz;
    while (zc > 8 && cols * ((g->nz + zc - 1) / zc) < 2048) zc = (zc + 1) / 2;
    a.zc = zc;
    const int nchunks = (g->nz + zc - 1) / zc;
    a.c1 = c1;
    a.U = g->face[0].p;
    a.V = g->face[1].p;
    a.W = g->face[2].p;
    a.lA0 = lA[0];
    a.lA1 = lA[1];
    a.lA2 = lA[2];
    a.l1 = l1;
    hipLaunchKernelGGL((k_adj_anti_t<TX, TY>), dim3((unsigned)(cols * nchunks)),
                       dim3(tile::GeoB<TX, TY>::NTH), 0, s, a);
    MPD_HIP(hipGetLastError());
    return MPD_OK;
}

// donor pass transposed: lin += l1 + gather over the Courant faces.  One thread per
// (x, y) column and z-chunk (workgroups of 64 x 4 columns, coalesced along x): the own
// column's l1 below / above and W below are carried from plane to plane, the in-plane
// neighbours are cached loads.  Per face (x, y, z): the upper face's outflow term, then
// the lower face's, in that order.
constexpr int D1X = 64, D1Y = 4;
__global__ void __launch_bounds__(D1X * D1Y)
k_adj_donor1(int nx, int ny, int nz, int zc, const double* __restrict__ U,
             const double* __restrict__ V, const double* __restrict__ W,
             const double* __restrict__ l1, double* __restrict__ lin)
{
    const int i = blockIdx.x * D1X + (int)threadIdx.x % D1X;
    const int j = blockIdx.y * D1Y + (int)threadIdx.x / D1X;
    if (i >= nx || j >= ny) return;
    const int z0 = blockIdx.z * zc, z1 = min(z0 + zc, nz);
    const size_t np = (size_t)nx * ny;
    const int il = i == 0 ? nx - 1 : i - 1, ir = i + 1 == nx ? 0 : i + 1;
    const int jd = j == 0 ? ny - 1 : j - 1, ju = j + 1 == ny ? 0 : j + 1;
    const size_t c = (size_t)j * nx + i;
    const size_t cxl = (size_t)j * nx + il, cxr = (size_t)j * nx + ir;
    const size_t cyd = (size_t)jd * nx + i, cyu = (size_t)ju * nx + i;
    const size_t below = (size_t)(z0 == 0 ? nz - 1 : z0 - 1) * np;
    double lm = l1[below + c], wm = W[below + c];
    size_t o = (size_t)z0 * np;
    double lc = l1[o + c];
    for (int k = z0; k < z1; k++, o += np) {
        const double lp = l1[(k + 1 == nz ? 0 : o + np) + c];
        double acc = lc;
        double v = U[o + c];
        acc += 0.5 * (v + ab(v)) * (l1[o + cxr] - lc);
        v = U[o + cxl];
        acc += 0.5 * (v - ab(v)) * (lc - l1[o + cxl]);
        v = V[o + c];
        acc += 0.5 * (v + ab(v)) * (l1[o + cyu] - lc);
        v = V[o + cyd];
        acc += 0.5 * (v - ab(v)) * (lc - l1[o + cyd]);
        const double wc = W[o + c];
        acc += 0.5 * (wc + ab(wc)) * (lp - lc);
        acc += 0.5 * (wm - ab(wm)) * (lc - lm);
        lin[o + c] += acc;
        lm = lc;
        lc = lp;
        wm = wc;
    }
}

// receptor residuals: lam[rec] += c[rec] - obs (and the loss terms), one thread each
__global__ void k_receptors(const double* __restrict__ c, const int* __restrict__ rec, int nrec,
                            const double* __restrict__ obs, double* __restrict__ res,
                            double* __restrict__ lam)
{
    int r = blockIdx.x * TPB + threadIdx.x;
    if (r >= nrec) return;
    double d = c[rec[r]] - obs[r];
    if (res) res[r] = d;
    if (lam) lam[rec[r]] += d;  // receptor cells are distinct (checked on the host)
}

// Forward intermediates of one step, stored by the fused forward kernels.
struct Fwd {
    DevBuf c1, A[3], bup, bdn;
    int* code = nullptr;
    Fwd() = default;
    Fwd(const Fwd&) = delete;
    Fwd& operator=(const Fwd&) = delete;
    ~Fwd()
    {
        if (code) (void)hipFree(code);
    }
    int alloc(size_t cells, bool nonosc)
    {
        int rc;
        DevBuf* all[] = {&c1, &A[0], &A[1], &A[2], &bup, &bdn};
        for (int i = 0; i < (nonosc ? 6 : 4); i++)
            if ((rc = all[i]->alloc(cells))) return rc;
        if (nonosc) MPD_HIP(hipMalloc((void**)&code, cells * sizeof(int)));
        return MPD_OK;
    }
    // one exact step cin -> cout storing this step's intermediates
    int step(mpd_grid_s* g, const double* cin, double* cout)
    {
        double* const A3[3] = {A[0].p, A[1].p, A[2].p};
        return fused_store_step_exact(g, cin, cout, c1.p, A3, bup.p, bdn.p, code);
    }
};

// Adjoint intermediates of one step.
struct Work {
    DevBuf lA[3], iu, id, lain, laout, l1;
    int alloc(size_t cells)
    {
        int rc;
        DevBuf* all[] = {&lA[0], &lA[1], &lA[2], &iu, &id, &lain, &laout, &l1};
        for (DevBuf* d : all)
            if ((rc = d->alloc(cells))) return rc;
        return MPD_OK;
    }
};

// Forward record of one checkpoint segment of K steps: the state entering each step and
// the step's intermediates.
struct Record {
    std::vector<DevBuf> st;
    std::vector<Fwd> fw;
    static size_t step_bytes(size_t cells, bool nonosc)
    {
        return cells * (nonosc ? 8 * 7 + 4 : 8 * 5);
    }
    int alloc(int K, size_t cells, bool nonosc)
    {
        st = std::vector<DevBuf>(K);
        fw = std::vector<Fwd>(K);
        int rc;
        for (int m = 0; m < K; m++)
            if ((rc = st[m].alloc(cells)) || (rc = fw[m].alloc(cells, nonosc))) return rc;
        return MPD_OK;
    }
};

// Device memory the gradient driver leaves free when it sizes the kept records by the
// free memory.
constexpr size_t kTapeReserve = (size_t)8 << 30;

// lin = A'(c_in)^T lout given the step's stored forward f; lq += lout (lq may be null)
int adjoint_step(mpd_grid_s* g, Work& w, const Fwd& f, const double* cin, const double* lout,
                 double* lin, double* lq)
{
    hipStream_t s = stream();
    Geo geo{g->nx, g->ny, g->nz, (size_t)g->nx * g->ny, g->ncell};
    const unsigned nb = nblk(g->ncell);
    const double *U = g->face[0].p, *V = g->face[1].p, *W = g->face[2].p;
    AdjBufs b{U, V, W, cin, f.c1.p, f.A[0].p, f.A[1].p, f.A[2].p, f.bup.p, f.bdn.p,
              f.code, g->nonosc};
    if (g->nonosc && g->adj_tiled) {
        const double* in[6] = {f.c1.p, f.A[0].p, f.A[1].p, f.A[2].p, f.bup.p, f.bdn.p};
        double* const lA[3] = {w.lA[0].p, w.lA[1].p, w.lA[2].p};
        int rc;
        if ((rc = lim_tiled(g, in, f.code, lout, w.l1.p, lA, lin, lq, s))) return rc;
    } else {
        hipLaunchKernelGGL(k_adj_donor_limit, dim3(nb), dim3(TPB), 0, s, geo, b, lout, w.l1.p,
                           w.lA[0].p, w.lA[1].p, w.lA[2].p, w.iu.p, w.id.p, w.lain.p,
                           w.laout.p, lq);
        if (g->nonosc)
            hipLaunchKernelGGL(k_adj_limit3, dim3(nb), dim3(TPB), 0, s, geo, b, w.iu.p, w.id.p,
                               w.lain.p, w.laout.p, lin, w.l1.p, w.lA[0].p, w.lA[1].p,
                               w.lA[2].p);
        else
            MPD_HIP(hipMemsetAsync(lin, 0, g->ncell * 8, s));
    }
    if (g->adj_scatter) {
        hipLaunchKernelGGL(k_adj_anti_atomic, dim3(nb), dim3(TPB), 0, s, geo, b, w.lA[0].p,
                           w.lA[1].p, w.lA[2].p, w.l1.p);
    } else {
        double* const lA[3] = {w.lA[0].p, w.lA[1].p, w.lA[2].p};
        int rc;
        if ((rc = anti_tiled(g, f.c1.p, lA, w.l1.p, s))) return rc;
    }
    {
        const unsigned gx = (g->nx + D1X - 1) / D1X, gy = (g->ny + D1Y - 1) / D1Y;
        int zc = g->nz;
        while (zc > 8 && (long long)gx * gy * ((g->nz + zc - 1) / zc) < 2048) zc = (zc + 1) / 2;
        const unsigned gz = (g->nz + zc - 1) / zc;
        hipLaunchKernelGGL(k_adj_donor1, dim3(gx, gy, gz), dim3(D1X * D1Y), 0, s, g->nx, g->ny,
                           g->nz, zc, U, V, W, (const double*)w.l1.p, lin);
    }
    MPD_HIP(hipGetLastError());
    return MPD_OK;
}

// The adjoint linearises the exact fused step; grids it cannot run on are refused.
int adjoint_supported(const mpd_grid_s* g)
{
    if (g->boundary != MPD_PERIODIC || g->iord != 2)
        return fail(MPD_ERR_UNSUPPORTED, "the adjoint covers PERIODIC grids with iord = 2");
    if (!fused_plane_fits(g->nx, g->ny))
        return fail(MPD_ERR_UNSUPPORTED, "the adjoint needs z-planes of at most 2^29 cells");
    if ((g->ny + D1Y - 1) / D1Y > 65535)
        return fail(MPD_ERR_UNSUPPORTED, "the adjoint needs ny <= %d", 65535 * D1Y);
    return MPD_OK;
}

}  // namespace

// One adjoint step on host arrays (for tests): lin = A'(c)^T lout, lq += lout.
extern "C" int mpd_set_adjoint_scatter(mpd_grid g, int atomic)
{
    MPD_API_BEGIN
    clear_error();
    if (!g) return fail(MPD_ERR_ARG, "mpd_set_adjoint_scatter: NULL grid");
    g->adj_scatter = atomic ? 1 : 0;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_set_adjoint_tiled(mpd_grid g, int tiled)
{
    MPD_API_BEGIN
    clear_error();
    if (!g) return fail(MPD_ERR_ARG, "mpd_set_adjoint_tiled: NULL grid");
    g->adj_tiled = tiled ? 1 : 0;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_set_adjoint_tape(mpd_grid g, long long max_bytes)
{
    MPD_API_BEGIN
    clear_error();
    if (!g) return fail(MPD_ERR_ARG, "mpd_set_adjoint_tape: NULL grid");
    g->adj_tape = max_bytes;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_get_adjoint_tape(mpd_grid g, long long* max_bytes, int* kept_segments)
{
    MPD_API_BEGIN
    clear_error();
    if (!g) return fail(MPD_ERR_ARG, "mpd_get_adjoint_tape: NULL grid");
    if (max_bytes) *max_bytes = g->adj_tape;
    if (kept_segments) *kept_segments = g->adj_kept;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_adjoint_step(mpd_grid g, const double* c, const double* lout, double* lin,
                                double* lq)
{
    MPD_API_BEGIN
    clear_error();
    if (!g || !c || !lout || !lin) return fail(MPD_ERR_ARG, "mpd_adjoint_step: NULL argument");
    int rc;
    if ((rc = adjoint_supported(g))) return rc;
    if (!g->has_winds) return fail(MPD_ERR_ARG, "mpd_adjoint_step: set the Courant numbers first");
    if ((rc = ensure_fused_scratch(g))) return rc;
    Work w;
    Fwd f;
    if ((rc = w.alloc(g->ncell)) || (rc = f.alloc(g->ncell, g->nonosc))) return rc;
    DevBuf dc, dl, di, dq, dout;
    hipStream_t s = stream();
    if ((rc = dc.alloc(g->ncell)) || (rc = dl.alloc(g->ncell)) || (rc = di.alloc(g->ncell)) ||
        (rc = dq.alloc(g->ncell)) || (rc = dout.alloc(g->ncell)))
        return rc;
    MPD_HIP(hipMemcpyAsync(dc.p, c, g->ncell * 8, hipMemcpyHostToDevice, s));
    MPD_HIP(hipMemcpyAsync(dl.p, lout, g->ncell * 8, hipMemcpyHostToDevice, s));
    if (lq) MPD_HIP(hipMemcpyAsync(dq.p, lq, g->ncell * 8, hipMemcpyHostToDevice, s));
    if ((rc = f.step(g, dc.p, dout.p))) return rc;
    if ((rc = adjoint_step(g, w, f, dc.p, dl.p, di.p, lq ? dq.p : nullptr))) return rc;
    MPD_HIP(hipMemcpyAsync(lin, di.p, g->ncell * 8, hipMemcpyDeviceToHost, s));
    if (lq) MPD_HIP(hipMemcpyAsync(lq, dq.p, g->ncell * 8, hipMemcpyDeviceToHost, s));
    MPD_HIP(hipStreamSynchronize(s));
    return MPD_OK;
    MPD_API_END
}

// Forward run from the grid's field recording the receptor values every sample_every
// steps: out[(s-1)*nrec + r] = c^{s*sample_every}[rec_r].  The grid ends at c^nsteps.
extern "C" int mpd_sample_receptors(mpd_grid g, int nsteps, int nrec, const int* receptors,
                                    int sample_every, double* out)
{
    MPD_API_BEGIN
    clear_error();
    if (!g || nsteps < 1 || nrec < 1 || !receptors || sample_every < 1 || !out)
        return fail(MPD_ERR_ARG, "mpd_sample_receptors: bad arguments");
    // the loop below steps with periodic_step(), whose kernels index the face arrays in
    // the periodic layout (ncell elements each); a FRAME grid's arrays are shorter
    if (g->boundary != MPD_PERIODIC)
        return fail(MPD_ERR_UNSUPPORTED, "mpd_sample_receptors: PERIODIC grids only");
    if (!g->has_winds || !g->has_field)
        return fail(MPD_ERR_ARG, "mpd_sample_receptors: set the Courant numbers and the field first");
    for (int q = 0; q < nrec; q++)
        if (receptors[q] < 0 || (size_t)receptors[q] >= g->ncell)
            return fail(MPD_ERR_ARG, "receptor index out of range");
    const int nsamp = nsteps / sample_every;
    hipStream_t s = stream();
    int* drec = nullptr;
    MPD_HIP(hipMalloc((void**)&drec, nrec * sizeof(int)));
    struct FreeI {
        int* p;
        ~FreeI() { if (p) (void)hipFree(p); }
    } freei{drec};
    MPD_HIP(hipMemcpy(drec, receptors, nrec * sizeof(int), hipMemcpyHostToDevice));
    DevBuf zero, vals;
    int rc;
    if ((rc = zero.alloc(nrec)) || (rc = vals.alloc((size_t)std::max(nsamp, 1) * nrec))) return rc;
    std::vector<double> z(nrec, 0.0);
    MPD_HIP(hipMemcpy(zero.p, z.data(), nrec * 8, hipMemcpyHostToDevice));
    for (int step = 0; step < nsteps; step++) {
        if ((rc = periodic_step(g))) return rc;
        if ((step + 1) % sample_every == 0) {
            int sm = (step + 1) / sample_every - 1;
            hipLaunchKernelGGL(k_receptors, dim3(nblk(nrec)), dim3(TPB), 0, s, g->c[g->cur].p,
                               drec, nrec, zero.p, vals.p + (size_t)sm * nrec, (double*)nullptr);
        }
    }
    if (nsamp)
        MPD_HIP(hipMemcpyAsync(out, vals.p, (size_t)nsamp * nrec * 8, hipMemcpyDeviceToHost, s));
    MPD_HIP(hipStreamSynchronize(s));
    return MPD_OK;
    MPD_API_END
}

// Receptor-misfit loss and its gradient.  From the grid's current field c^0 and
// emission q (none if unset): c^{n+1} = A(c^n) + q for n < nsteps;
// J = 1/2 sum over samples s = 1 .. nsteps/sample_every of
//     sum_r (c^{s*sample_every}[rec_r] - obs[(s-1)*nrec + r])^2.
// The forward keeps c^n every `checkpoint_every` steps; the backward sweep recomputes
// each segment from its checkpoint and runs the adjoint steps in reverse.  grad_c0 /
// grad_q (host, n cells each) may be NULL; with both NULL only J is computed.  On
// return the grid holds c^nsteps.  times (optional, 3 doubles): forward and backward
// milliseconds, and the part of the backward spent recomputing checkpoint segments
// (HIP events).
extern "C" int mpd_adjoint_run(mpd_grid g, int nsteps, int checkpoint_every, int nrec,
                               const int* receptors, int sample_every, const double* obs,
                               double* J, double* grad_c0, double* grad_q, double* times)
{
    MPD_API_BEGIN
    clear_error();
    if (!g || nsteps < 1 || checkpoint_every < 1 || nrec < 1 || !receptors ||
        sample_every < 1 || !obs || !J)
        return fail(MPD_ERR_ARG, "mpd_adjoint_run: bad arguments");
    int rc;
    if ((rc = adjoint_supported(g))) return rc;
    if (!g->has_winds || !g->has_field)
        return fail(MPD_ERR_ARG, "mpd_adjoint_run: set the Courant numbers and the field first");
    {
        std::vector<int> r(receptors, receptors + nrec);
        std::sort(r.begin(), r.end());
        for (int q = 0; q < nrec; q++)
            if (r[q] < 0 || (size_t)r[q] >= g->ncell || (q && r[q] == r[q - 1]))
                return fail(MPD_ERR_ARG, "receptors must be distinct cell indices in range");
    }
    const bool want_grad = grad_c0 || grad_q;
    const int nsamp = nsteps / sample_every;
    const int K = checkpoint_every;
    const int nck = (nsteps + K - 1) / K;  // checkpoints at 0, K, 2K, ...
    hipStream_t s = stream();
    const size_t n = g->ncell;
    // arithmetic of the forward: the adjoint linearises exact-arithmetic steps
    const int saved_arith = g->arith;
    g->arith = MPD_ARITH_EXACT;
    struct Restore {
        mpd_grid_s* g;
        int a;
        ~Restore() { g->arith = a; }
    } restore{g, saved_arith};

    DevBuf drec_b, dobs, dres;
    int* drec = nullptr;
    MPD_HIP(hipMalloc((void**)&drec, nrec * sizeof(int)));
    struct FreeI {
        int* p;
        ~FreeI() { if (p) (void)hipFree(p); }
    } freei{drec};
    MPD_HIP(hipMemcpy(drec, receptors, nrec * sizeof(int), hipMemcpyHostToDevice));
    if ((rc = dobs.alloc((size_t)std::max(nsamp, 1) * nrec)) || (rc = dres.alloc((size_t)std::max(nsamp, 1) * nrec)))
        return rc;
    if (nsamp) MPD_HIP(hipMemcpy(dobs.p, obs, (size_t)nsamp * nrec * 8, hipMemcpyHostToDevice));
    // Storage of the backward sweep, allocated before the timed sweeps: the adjoint
    // intermediates and the forward records of whole segments (each step's state and
    // intermediates, about ten arrays per step).  Record 0 is the working record: it keeps
    // the last segment from the forward sweep and then takes each recomputed segment.
    // Records 1 .. keep-1 keep the segments before the last one from the forward sweep,
    // as many as the tape budget and the free device memory allow; only the segments
    // before those are recomputed from their checkpoints.
    Work w;
    DevBuf lam, lnext, gq, tmp, final_state;
    std::vector<std::unique_ptr<Record>> recs;
    int keep = 0;
    if (want_grad) {
        if ((rc = w.alloc(n)) || (rc = lam.alloc(n)) || (rc = lnext.alloc(n)) ||
            (rc = gq.alloc(n)) || (rc = tmp.alloc(n)) || (rc = final_state.alloc(n)))
            return rc;
        const size_t rec_bytes = (size_t)K * Record::step_bytes(n, g->nonosc);
        int extra = 0;
        if (g->adj_tape != 0 && nck > 1) {
            size_t freeb = 0, totalb = 0;
            MPD_HIP(hipMemGetInfo(&freeb, &totalb));
            // the working record and the checkpoints of the recomputed segments are still
            // to be allocated; a reserve stays free
            const size_t need = rec_bytes + (size_t)nck * n * 8 + kTapeReserve;
            size_t fit = freeb > need ? (freeb - need) / rec_bytes : 0;
            if (g->adj_tape > 0) fit = std::min(fit, (size_t)g->adj_tape / rec_bytes);
            extra = (int)std::min(fit, (size_t)(nck - 1));
        }
        keep = 1 + extra;
        for (int r = 0; r < keep; r++) {
            recs.emplace_back(new Record);
            if ((rc = recs.back()->alloc(K, n, g->nonosc)))
                return rc == MPD_ERR_NOMEM
                           ? fail(rc, "mpd_adjoint_run: device memory for %d stored steps of "
                                      "%zu cells exhausted; use a shorter checkpoint interval "
                                      "or a smaller tape budget",
                                  K * (r + 1), n)
                           : rc;
        }
    }
    // segments [first_kept, nck) keep their records from the forward sweep; segment c
    // of those is held by record nck - 1 - c
    const int first_kept = nck - keep;
    g->adj_kept = keep;
    std::vector<DevBuf> ck(want_grad ? first_kept : 0);
    for (auto& b : ck)
        if ((rc = b.alloc(n))) return rc;
    // c^{cK}: a checkpoint, or the first state of a kept record
    auto ckpt = [&](int c) -> const double* {
        return c < first_kept ? ck[c].p : recs[nck - 1 - c]->st[0].p;
    };

    hipEvent_t e0, e1, e1b, e2;
    MPD_HIP(hipEventCreate(&e0));
    MPD_HIP(hipEventCreate(&e1));
    MPD_HIP(hipEventCreate(&e1b));
    MPD_HIP(hipEventCreate(&e2));
    MPD_HIP(hipEventRecord(e0, s));
    // forward sweep: plain steps on the grid's buffers up to the first kept segment, then
    // steps that store their record
    for (int step = 0; step < nsteps; step++) {
        const int c = step / K, m = step - c * K;
        const double* out;
        if (c < first_kept) {
            if (want_grad && m == 0)
                MPD_HIP(hipMemcpyAsync(ck[c].p, g->c[g->cur].p, n * 8, hipMemcpyDeviceToDevice, s));
            if ((rc = periodic_step(g))) return rc;
            out = g->c[g->cur].p;
        } else {
            Record& R = *recs[nck - 1 - c];
            if (step == first_kept * K)
                MPD_HIP(hipMemcpyAsync(R.st[0].p, g->c[g->cur].p, n * 8, hipMemcpyDeviceToDevice, s));
            const int len = std::min(nsteps, (c + 1) * K) - c * K;
            double* dst = m + 1 < len ? R.st[m + 1].p
                                      : (c + 1 < nck ? recs[nck - 2 - c]->st[0].p : final_state.p);
            if ((rc = R.fw[m].step(g, R.st[m].p, dst))) return rc;
            out = dst;
        }
        if ((step + 1) % sample_every == 0) {
            int sm = (step + 1) / sample_every - 1;
            hipLaunchKernelGGL(k_receptors, dim3(nblk(nrec)), dim3(TPB), 0, s, out, drec, nrec,
                               dobs.p + (size_t)sm * nrec, dres.p + (size_t)sm * nrec,
                               (double*)nullptr);
        }
    }
    MPD_HIP(hipEventRecord(e1, s));
    std::vector<double> res((size_t)std::max(nsamp, 1) * nrec, 0.0);
    if (nsamp) MPD_HIP(hipMemcpyAsync(res.data(), dres.p, (size_t)nsamp * nrec * 8, hipMemcpyDeviceToHost, s));
    MPD_HIP(hipStreamSynchronize(s));
    double jj = 0.0;
    for (size_t q = 0; q < (size_t)nsamp * nrec; q++) jj += 0.5 * res[q] * res[q];
    *J = jj;

    // start / end of each segment's recompute
    std::vector<hipEvent_t> rec_ev(want_grad ? 2 * nck : 0, nullptr);
    struct FreeEv {
        std::vector<hipEvent_t>& v;
        ~FreeEv()
        {
            for (hipEvent_t e : v)
                if (e) (void)hipEventDestroy(e);
        }
    } free_ev{rec_ev};
    for (hipEvent_t& e : rec_ev) MPD_HIP(hipEventCreate(&e));
    MPD_HIP(hipEventRecord(e1b, s));
    if (want_grad) {
        MPD_HIP(hipMemsetAsync(lam.p, 0, n * 8, s));
        MPD_HIP(hipMemsetAsync(gq.p, 0, n * 8, s));
        for (int c = nck - 1; c >= 0; c--) {
            const int s0 = c * K, s1 = std::min(nsteps, s0 + K);
            // the residual of c^{s1}, the next segment's first state (or the final field),
            // is added before the recompute below overwrites the working record
            if (s1 % sample_every == 0) {
                const double* cs = c + 1 < nck ? ckpt(c + 1) : final_state.p;
                int sm = s1 / sample_every - 1;
                hipLaunchKernelGGL(k_receptors, dim3(nblk(nrec)), dim3(TPB), 0, s, cs, drec,
                                   nrec, dobs.p + (size_t)sm * nrec, (double*)nullptr, lam.p);
            }
            Record* R;
            if (c >= first_kept) {
                R = recs[nck - 1 - c].get();
            } else {
                // recompute c^{s0} .. c^{s1-1} into the working record with each step's
                // intermediates (the last step's result, c^{s1}, is not needed again)
                R = recs[0].get();
                MPD_HIP(hipEventRecord(rec_ev[2 * c], s));
                MPD_HIP(hipMemcpyAsync(R->st[0].p, ck[c].p, n * 8, hipMemcpyDeviceToDevice, s));
                for (int m = 0; m < s1 - s0; m++)
                    if ((rc = R->fw[m].step(g, R->st[m].p, m + 1 < s1 - s0 ? R->st[m + 1].p : tmp.p)))
                        return rc;
                MPD_HIP(hipEventRecord(rec_ev[2 * c + 1], s));
            }
            for (int step = s1; step > s0; step--) {  // lam holds lam^{step}
                if (step < s1 && step % sample_every == 0) {
                    int sm = step / sample_every - 1;
                    hipLaunchKernelGGL(k_receptors, dim3(nblk(nrec)), dim3(TPB), 0, s,
                                       (const double*)R->st[step - s0].p, drec, nrec,
                                       dobs.p + (size_t)sm * nrec, (double*)nullptr, lam.p);
                }
                if ((rc = adjoint_step(g, w, R->fw[step - 1 - s0], R->st[step - 1 - s0].p, lam.p,
                                       lnext.p, gq.p)))
                    return rc;
                std::swap(lam.p, lnext.p);
            }
        }
        MPD_HIP(hipEventRecord(e2, s));
        MPD_HIP(hipMemcpyAsync(g->c[g->cur].p, final_state.p, n * 8, hipMemcpyDeviceToDevice, s));
        if (grad_c0) MPD_HIP(hipMemcpyAsync(grad_c0, lam.p, n * 8, hipMemcpyDeviceToHost, s));
        if (grad_q) MPD_HIP(hipMemcpyAsync(grad_q, gq.p, n * 8, hipMemcpyDeviceToHost, s));
        MPD_HIP(hipStreamSynchronize(s));
    } else {
        MPD_HIP(hipEventRecord(e2, s));
        MPD_HIP(hipStreamSynchronize(s));
    }
    float f1 = 0.f, f2 = 0.f;
    (void)hipEventElapsedTime(&f1, e0, e1);
    (void)hipEventElapsedTime(&f2, e1b, e2);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e1b);
    (void)hipEventDestroy(e2);
    if (times) {
        times[0] = f1;
        times[1] = f2;
        double fr = 0.0;  // (events recorded for the recomputed segments only)
        for (size_t q = 0; q < 2 * (size_t)std::max(first_kept, 0) && q + 1 < rec_ev.size(); q += 2) {
            float f = 0.f;
            (void)hipEventElapsedTime(&f, rec_ev[q], rec_ev[q + 1]);
            fr += f;
        }
        times[2] = fr;
    }
    return MPD_OK;
    MPD_API_END
}
