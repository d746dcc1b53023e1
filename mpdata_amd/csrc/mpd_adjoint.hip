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
#include <vector>

#include "mpd_formulas.hpp"
#include "mpd_fused.hpp"
#include "mpd_grid.hpp"

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
    const double *U, *V, *W, *c0, *c1, *A0, *A1, *A2, *bup, *bdn, *ain, *aout;
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

// corrective donor pass transposed: l1 = lout + gather; lAl of the upper faces; lq += lout
__global__ void k_adj_donor2(Geo g, AdjBufs b, const double* __restrict__ lout,
                             double* __restrict__ l1, double* __restrict__ lAl0,
                             double* __restrict__ lAl1, double* __restrict__ lAl2,
                             double* __restrict__ lq)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= g.n) return;
    size_t lo[3], up[3];
    nbrs(g, t, lo, up);
    double* lAl[3] = {lAl0, lAl1, lAl2};
    double acc = lout[t];
    for (int d = 0; d < 3; d++) {
        double gl, gr, gv;
        // upper face (t, up): lam_f = lout[up] - lout[t].  The ratio arrays exist only
        // with the limiter: read them only then.
        double au = fam_a(b, d, t);
        if (b.nonosc) au = limited(au, b.bup[t], b.bdn[t], b.bup[up[d]], b.bdn[up[d]]);
        fpart(b.c1[t], b.c1[up[d]], au, gl, gr, gv);
        double lf = lout[up[d]] - lout[t];
        acc += gl * lf;
        lAl[d][t] = gv * lf;
        // lower face (lo, t)
        double al = fam_a(b, d, lo[d]);
        if (b.nonosc) al = limited(al, b.bup[lo[d]], b.bdn[lo[d]], b.bup[t], b.bdn[t]);
        fpart(b.c1[lo[d]], b.c1[t], al, gl, gr, gv);
        acc += gr * (lout[t] - lout[lo[d]]);
    }
    l1[t] = acc;
    if (lq) lq[t] += lout[t];
}

// limiter transposed, part 1: lA = m * lAl on the upper faces; the ratio adjoints,
// gathered from the six faces of the cell, turned into bound and flux-sum adjoints
__global__ void k_adj_limit1(Geo g, AdjBufs b, const double* __restrict__ lAl0,
                             const double* __restrict__ lAl1, const double* __restrict__ lAl2,
                             double* __restrict__ lA0, double* __restrict__ lA1,
                             double* __restrict__ lA2, double* __restrict__ iu,
                             double* __restrict__ id, double* __restrict__ lain,
                             double* __restrict__ laout)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= g.n) return;
    size_t lo[3], up[3];
    nbrs(g, t, lo, up);
    const double* lAl[3] = {lAl0, lAl1, lAl2};
    double* lA[3] = {lA0, lA1, lA2};
    double su = 0.0, sd = 0.0;
    for (int d = 0; d < 3; d++) {
        // upper face: t is the lower cell
        size_t r = up[d];
        double a = fam_a(b, d, t), L = lAl[d][t], m = 1.0;
        if (a > 0) {
            int w = pick(b.bdn[t], b.bup[r]);
            m = w == 0 ? 1.0 : (w == 1 ? b.bdn[t] : b.bup[r]);
            if (w == 1) sd += a * L;
        } else if (a < 0) {
            int w = pick(b.bup[t], b.bdn[r]);
            m = w == 0 ? 1.0 : (w == 1 ? b.bup[t] : b.bdn[r]);
            if (w == 1) su += a * L;
        } else {
            m = 0.0;
        }
        lA[d][t] = m * L;
        // lower face: t is the upper cell
        size_t l = lo[d];
        a = fam_a(b, d, l);
        L = lAl[d][l];
        if (a > 0) {
            if (pick(b.bdn[l], b.bup[t]) == 2) su += a * L;
        } else if (a < 0) {
            if (pick(b.bup[l], b.bdn[t]) == 2) sd += a * L;
        }
    }
    // ratio adjoints -> bound / flux-sum adjoints (pointwise).  A ratio that was not
    // selected carries no adjoint and may be inf: skipped.
    double u = 0.0, dn = 0.0, li = 0.0, lo2 = 0.0;
    if (su != 0.0) {
        u = su / (b.ain[t] + MINV);
        li = -u * b.bup[t];
    }
    if (sd != 0.0) {
        dn = sd / (b.aout[t] + MINV);
        lo2 = -dn * b.bdn[t];
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

// antidiffusive velocities transposed, part 1: per face the coefficients of its ten
// cells times lam_A: kp, km, ku1, kd1, ku2, kd2 (coef[fam*6 + role] arrays)
__global__ void k_adj_anti1(Geo g, AdjBufs b, const double* __restrict__ lA0,
                            const double* __restrict__ lA1, const double* __restrict__ lA2,
                            double* const* __restrict__ coef)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= g.n) return;
    int i, j, k;
    g.ijk(t, i, j, k);
    const double* lA[3] = {lA0, lA1, lA2};
    const double* c = b.c1;
    for (int fam = 0; fam < 3; fam++) {
        FaceSt f = face_st(g, b.U, b.V, b.W, i, j, k, fam);
        double L = lA[fam][t];
        double p = c[f.p], m = c[f.m];
        double K = ab(f.a) * (1.0 - ab(f.a));
        double den = p + m + MINV;
        double n1 = c[f.up1[0]] + c[f.up1[1]] - c[f.dn1[0]] - c[f.dn1[1]];
        double D1 = c[f.up1[0]] + c[f.up1[1]] + c[f.dn1[0]] + c[f.dn1[1]] + MINV;
        double n2 = c[f.up2[0]] + c[f.up2[1]] - c[f.dn2[0]] - c[f.dn2[1]];
        double D2 = c[f.up2[0]] + c[f.up2[1]] + c[f.dn2[0]] + c[f.dn2[1]] + MINV;
        double gb = -0.5 * f.a;
        double* o = nullptr;
        o = coef[fam * 6 + 0]; o[t] = L * (K * (((2.0 * m + MINV) / den) / den));
        o = coef[fam * 6 + 1]; o[t] = L * (-K * (((2.0 * p + MINV) / den) / den));
        o = coef[fam * 6 + 2]; o[t] = L * (gb * f.s1 * (((D1 - n1) / D1) / D1));
        o = coef[fam * 6 + 3]; o[t] = L * (gb * f.s1 * (((-D1 - n1) / D1) / D1));
        o = coef[fam * 6 + 4]; o[t] = L * (gb * f.s2 * (((D2 - n2) / D2) / D2));
        o = coef[fam * 6 + 5]; o[t] = L * (gb * f.s2 * (((-D2 - n2) / D2) / D2));
    }
}

// Scatter variant of parts 1 + 2: each face adds its ten contributions to l1 with
// hardware FP64 atomics (no coefficient arrays; summation order, and so the last bits,
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

// part 2: every cell gathers the coefficients of the faces whose stencil contains it
// (the stencils of face_st read backwards)
__global__ void k_adj_anti2(Geo g, double* const* __restrict__ coef, double* __restrict__ l1)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= g.n) return;
    int i, j, k;
    g.ijk(t, i, j, k);
    const double *const *C = coef;
    auto K = [&](int fam, int role, int di, int dj, int dk) {
        return C[fam * 6 + role][g.at(i + di, j + dj, k + dk)];
    };
    double s = 0.0;
    // x-faces (i+1/2): p=(i+1), m=(i); up1 (i+1,j+1),(i,j+1); dn1 (i,j-1),(i+1,j-1);
    // up2 (i+1,k+1),(i,k+1); dn2 (i,k-1),(i+1,k-1)
    s += K(0, 0, -1, 0, 0) + K(0, 1, 0, 0, 0);
    s += K(0, 2, -1, -1, 0) + K(0, 2, 0, -1, 0);
    s += K(0, 3, 0, 1, 0) + K(0, 3, -1, 1, 0);
    s += K(0, 4, -1, 0, -1) + K(0, 4, 0, 0, -1);
    s += K(0, 5, 0, 0, 1) + K(0, 5, -1, 0, 1);
    // y-faces (j+1/2): p=(j+1), m=(j); up1 (i+1,j+1),(i+1,j); dn1 (i-1,j),(i-1,j+1);
    // up2 (j+1,k+1),(j,k+1); dn2 (j,k-1),(j+1,k-1)
    s += K(1, 0, 0, -1, 0) + K(1, 1, 0, 0, 0);
    s += K(1, 2, -1, -1, 0) + K(1, 2, -1, 0, 0);
    s += K(1, 3, 1, 0, 0) + K(1, 3, 1, -1, 0);
    s += K(1, 4, 0, -1, -1) + K(1, 4, 0, 0, -1);
    s += K(1, 5, 0, 0, 1) + K(1, 5, 0, -1, 1);
    // z-faces (k+1/2): p=(k+1), m=(k); up1 (i+1,k+1),(i+1,k); dn1 (i-1,k),(i-1,k+1);
    // up2 (j+1,k+1),(j+1,k); dn2 (j-1,k),(j-1,k+1)
    s += K(2, 0, 0, 0, -1) + K(2, 1, 0, 0, 0);
    s += K(2, 2, -1, 0, -1) + K(2, 2, -1, 0, 0);
    s += K(2, 3, 1, 0, 0) + K(2, 3, 1, 0, -1);
    s += K(2, 4, 0, -1, -1) + K(2, 4, 0, -1, 0);
    s += K(2, 5, 0, 1, 0) + K(2, 5, 0, 1, -1);
    l1[t] += s;
}

// donor pass transposed: lin += l1 + gather over the Courant faces
__global__ void k_adj_donor1(Geo g, AdjBufs b, const double* __restrict__ l1,
                             double* __restrict__ lin)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= g.n) return;
    size_t lo[3], up[3];
    nbrs(g, t, lo, up);
    const double* F[3] = {b.U, b.V, b.W};
    double acc = l1[t];
    for (int d = 0; d < 3; d++) {
        double v = F[d][t];
        acc += 0.5 * (v + ab(v)) * (l1[up[d]] - l1[t]);
        v = F[d][lo[d]];
        acc += 0.5 * (v - ab(v)) * (l1[t] - l1[lo[d]]);
    }
    lin[t] += acc;
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
    DevBuf c1, A[3], bup, bdn, ain, aout;
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
        DevBuf* all[] = {&c1, &A[0], &A[1], &A[2], &bup, &bdn, &ain, &aout};
        for (int i = 0; i < (nonosc ? 8 : 4); i++)
            if ((rc = all[i]->alloc(cells))) return rc;
        if (nonosc) MPD_HIP(hipMalloc((void**)&code, cells * sizeof(int)));
        return MPD_OK;
    }
    // one exact step cin -> cout storing this step's intermediates
    int step(mpd_grid_s* g, const double* cin, double* cout)
    {
        double* const A3[3] = {A[0].p, A[1].p, A[2].p};
        return fused_store_step_exact(g, cin, cout, c1.p, A3, bup.p, bdn.p, ain.p, aout.p, code);
    }
};

// Adjoint intermediates of one step.
struct Work {
    DevBuf lAl[3], lA[3], iu, id, lain, laout, l1;
    DevBuf coef[18];
    double** dcoef = nullptr;
    ~Work()
    {
        if (dcoef) (void)hipFree(dcoef);
    }
    int alloc(size_t cells)
    {
        int rc;
        DevBuf* all[] = {&lAl[0], &lAl[1], &lAl[2], &lA[0], &lA[1], &lA[2], &iu, &id, &lain,
                         &laout, &l1};
        for (DevBuf* d : all)
            if ((rc = d->alloc(cells))) return rc;
        for (auto& c : coef)
            if ((rc = c.alloc(cells))) return rc;
        MPD_HIP(hipMalloc((void**)&dcoef, 18 * sizeof(double*)));
        double* h[18];
        for (int q = 0; q < 18; q++) h[q] = coef[q].p;
        MPD_HIP(hipMemcpy(dcoef, h, sizeof h, hipMemcpyHostToDevice));
        return MPD_OK;
    }
};

// lin = A'(c_in)^T lout given the step's stored forward f; lq += lout (lq may be null)
int adjoint_step(mpd_grid_s* g, Work& w, const Fwd& f, const double* cin, const double* lout,
                 double* lin, double* lq)
{
    hipStream_t s = stream();
    Geo geo{g->nx, g->ny, g->nz, (size_t)g->nx * g->ny, g->ncell};
    const unsigned nb = nblk(g->ncell);
    const double *U = g->face[0].p, *V = g->face[1].p, *W = g->face[2].p;
    AdjBufs b{U, V, W, cin, f.c1.p, f.A[0].p, f.A[1].p, f.A[2].p, f.bup.p, f.bdn.p,
              f.ain.p, f.aout.p, f.code, g->nonosc};
    hipLaunchKernelGGL(k_adj_donor2, dim3(nb), dim3(TPB), 0, s, geo, b, lout, w.l1.p, w.lAl[0].p,
                       w.lAl[1].p, w.lAl[2].p, lq);
    if (g->nonosc) {
        hipLaunchKernelGGL(k_adj_limit1, dim3(nb), dim3(TPB), 0, s, geo, b, w.lAl[0].p,
                           w.lAl[1].p, w.lAl[2].p, w.lA[0].p, w.lA[1].p, w.lA[2].p, w.iu.p,
                           w.id.p, w.lain.p, w.laout.p);
        hipLaunchKernelGGL(k_adj_limit3, dim3(nb), dim3(TPB), 0, s, geo, b, w.iu.p, w.id.p,
                           w.lain.p, w.laout.p, lin, w.l1.p, w.lA[0].p, w.lA[1].p, w.lA[2].p);
    } else {
        for (int d = 0; d < 3; d++)
            MPD_HIP(hipMemcpyAsync(w.lA[d].p, w.lAl[d].p, g->ncell * 8, hipMemcpyDeviceToDevice, s));
        MPD_HIP(hipMemsetAsync(lin, 0, g->ncell * 8, s));
    }
    if (g->adj_scatter) {
        hipLaunchKernelGGL(k_adj_anti_atomic, dim3(nb), dim3(TPB), 0, s, geo, b, w.lA[0].p,
                           w.lA[1].p, w.lA[2].p, w.l1.p);
    } else {
        hipLaunchKernelGGL(k_adj_anti1, dim3(nb), dim3(TPB), 0, s, geo, b, w.lA[0].p,
                           w.lA[1].p, w.lA[2].p, w.dcoef);
        hipLaunchKernelGGL(k_adj_anti2, dim3(nb), dim3(TPB), 0, s, geo, w.dcoef, w.l1.p);
    }
    hipLaunchKernelGGL(k_adj_donor1, dim3(nb), dim3(TPB), 0, s, geo, b, w.l1.p, lin);
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
    return MPD_OK;
}

}  // namespace

// One adjoint step on host arrays (for tests): lin = A'(c)^T lout, lq += lout.
extern "C" int mpd_set_adjoint_scatter(mpd_grid g, int atomic)
{
    clear_error();
    if (!g) return fail(MPD_ERR_ARG, "mpd_set_adjoint_scatter: NULL grid");
    g->adj_scatter = atomic ? 1 : 0;
    return MPD_OK;
}

extern "C" int mpd_adjoint_step(mpd_grid g, const double* c, const double* lout, double* lin,
                                double* lq)
{
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
}

// Forward run from the grid's field recording the receptor values every sample_every
// steps: out[(s-1)*nrec + r] = c^{s*sample_every}[rec_r].  The grid ends at c^nsteps.
extern "C" int mpd_sample_receptors(mpd_grid g, int nsteps, int nrec, const int* receptors,
                                    int sample_every, double* out)
{
    clear_error();
    if (!g || nsteps < 1 || nrec < 1 || !receptors || sample_every < 1 || !out)
        return fail(MPD_ERR_ARG, "mpd_sample_receptors: bad arguments");
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
}

// Receptor-misfit loss and its gradient.  From the grid's current field c^0 and
// emission q (none if unset): c^{n+1} = A(c^n) + q for n < nsteps;
// J = 1/2 sum over samples s = 1 .. nsteps/sample_every of
//     sum_r (c^{s*sample_every}[rec_r] - obs[(s-1)*nrec + r])^2.
// The forward keeps c^n every `checkpoint_every` steps; the backward sweep recomputes
// each segment from its checkpoint and runs the adjoint steps in reverse.  grad_c0 /
// grad_q (host, n cells each) may be NULL; with both NULL only J is computed.  On
// return the grid holds c^nsteps.  times (optional, 2 doubles): forward and backward
// milliseconds (HIP events).
extern "C" int mpd_adjoint_run(mpd_grid g, int nsteps, int checkpoint_every, int nrec,
                               const int* receptors, int sample_every, const double* obs,
                               double* J, double* grad_c0, double* grad_q, double* times)
{
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
    std::vector<DevBuf> ck(want_grad ? nck : 0);
    for (auto& b : ck)
        if ((rc = b.alloc(n))) return rc;

    hipEvent_t e0, e1, e2;
    MPD_HIP(hipEventCreate(&e0));
    MPD_HIP(hipEventCreate(&e1));
    MPD_HIP(hipEventCreate(&e2));
    MPD_HIP(hipEventRecord(e0, s));
    // forward sweep
    for (int step = 0; step < nsteps; step++) {
        if (want_grad && step % K == 0)
            MPD_HIP(hipMemcpyAsync(ck[step / K].p, g->c[g->cur].p, n * 8, hipMemcpyDeviceToDevice, s));
        if ((rc = periodic_step(g))) return rc;
        if ((step + 1) % sample_every == 0) {
            int sm = (step + 1) / sample_every - 1;
            hipLaunchKernelGGL(k_receptors, dim3(nblk(nrec)), dim3(TPB), 0, s, g->c[g->cur].p,
                               drec, nrec, dobs.p + (size_t)sm * nrec, dres.p + (size_t)sm * nrec,
                               (double*)nullptr);
        }
    }
    MPD_HIP(hipEventRecord(e1, s));
    DevBuf final_state;
    if (want_grad) {  // c^nsteps: the last residual and the field handed back
        if ((rc = final_state.alloc(n))) return rc;
        MPD_HIP(hipMemcpyAsync(final_state.p, g->c[g->cur].p, n * 8, hipMemcpyDeviceToDevice, s));
    }
    std::vector<double> res((size_t)std::max(nsamp, 1) * nrec, 0.0);
    if (nsamp) MPD_HIP(hipMemcpyAsync(res.data(), dres.p, (size_t)nsamp * nrec * 8, hipMemcpyDeviceToHost, s));
    MPD_HIP(hipStreamSynchronize(s));
    double jj = 0.0;
    for (size_t q = 0; q < (size_t)nsamp * nrec; q++) jj += 0.5 * res[q] * res[q];
    *J = jj;

    if (want_grad) {
        Work w;
        if ((rc = w.alloc(n))) return rc;
        DevBuf lam, lnext, gq, tmp;
        if ((rc = lam.alloc(n)) || (rc = lnext.alloc(n)) || (rc = gq.alloc(n)) || (rc = tmp.alloc(n)))
            return rc;
        // a segment's states and the intermediates of its steps (about 10 arrays per step)
        std::vector<DevBuf> seg(K);
        std::vector<Fwd> fw(K);
        for (int m = 0; m < K; m++)
            if ((rc = seg[m].alloc(n)) || (rc = fw[m].alloc(n, g->nonosc)))
                return rc == MPD_ERR_NOMEM
                           ? fail(rc, "mpd_adjoint_run: device memory for %d stored steps of "
                                      "%zu cells exhausted; use a shorter checkpoint interval",
                                  K, n)
                           : rc;
        MPD_HIP(hipMemsetAsync(lam.p, 0, n * 8, s));
        MPD_HIP(hipMemsetAsync(gq.p, 0, n * 8, s));
        // a scratch grid view that steps the segment states: reuse g's buffers
        const int cur0 = g->cur;
        for (int c = nck - 1; c >= 0; c--) {
            const int s0 = c * K, s1 = std::min(nsteps, s0 + K);
            // recompute c^{s0} .. c^{s1-1} into seg[0 .. s1-s0-1] with each step's
            // intermediates (the last step's result, c^{s1}, is not needed again)
            MPD_HIP(hipMemcpyAsync(seg[0].p, ck[c].p, n * 8, hipMemcpyDeviceToDevice, s));
            for (int m = 0; m < s1 - s0; m++)
                if ((rc = fw[m].step(g, seg[m].p, m + 1 < s1 - s0 ? seg[m + 1].p : tmp.p)))
                    return rc;
            for (int step = s1; step > s0; step--) {  // lam holds lam^{step}
                if (step % sample_every == 0) {
                    // the residual needs c^{step}: c^{s1} is the next segment's first state
                    // (or the final field); inside the segment it is seg[step - s0]
                    const double* cs;
                    if (step == s1)
                        cs = (c + 1 < nck) ? ck[c + 1].p : nullptr;
                    else
                        cs = seg[step - s0].p;
                    if (!cs) cs = final_state.p;
                    int sm = step / sample_every - 1;
                    hipLaunchKernelGGL(k_receptors, dim3(nblk(nrec)), dim3(TPB), 0, s, cs, drec,
                                       nrec, dobs.p + (size_t)sm * nrec, (double*)nullptr, lam.p);
                }
                if ((rc = adjoint_step(g, w, fw[step - 1 - s0], seg[step - 1 - s0].p, lam.p,
                                       lnext.p, gq.p)))
                    return rc;
                std::swap(lam.p, lnext.p);
            }
        }
        (void)cur0;
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
    (void)hipEventElapsedTime(&f2, e1, e2);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (times) {
        times[0] = f1;
        times[1] = f2;
    }
    return MPD_OK;
}
