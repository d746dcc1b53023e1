/* CPU oracle for the derivatives of the periodic MPDATA step -- TEST INFRASTRUCTURE ONLY.
 *
 * No counterpart in the reference (it has no TLM or adjoint).  For one periodic step
 *     c_out = A(c_in) + q,   A = donor pass then one corrective pass (iord = 2),
 *                            optionally nonoscillatory (cpu_ref.c),
 * with the face Courant numbers U,V,W held fixed, this file gives
 *   orc_tlm_step  the tangent-linear model: dc_out = A'(c_in) dc_in + dq
 *   orc_adj_step  its transpose:            lam_in = A'(c_in)^T lam_out, lam_q += lam_out
 * Both differentiate the forward's own branch choices, so at a kink (a max/min tie,
 * a zero flux) they follow the branch the forward took: the TLM and the adjoint are
 * exact transposes of each other everywhere (dot-product test), and equal the true
 * derivative wherever the step is differentiable (finite-difference test).
 * Branch conventions (cpu_ref.c): mx2(a,b) = a > b ? a : b, mn2(a,b) = a < b ? a : b,
 * pos(a) = a > 0 ? a : 0, neg(a) = a < 0 ? a : 0, |v| = v > 0 ? v : -v.
 */
#include <float.h>
#include <stdlib.h>
#include <string.h>

#include "cpu_ref.h"

#define MINV DBL_MIN
#define PI(i, j, k) ((size_t)(k) * nx * ny + (size_t)(j) * nx + (i))

static inline double ab(double a) { return a > 0 ? a : -a; }
static inline double sg(double a) { return a > 0 ? 1.0 : -1.0; } /* d|a|/da of ab() */
static inline double F2(double l, double r, double v)
{
    double av = ab(v);
    return l * (v + av) + r * (v - av);
}

typedef struct {
    int nx, ny, nz;
    size_t n;
    int *xm, *xp, *ym, *yp, *zm, *zp; /* wrapped neighbour tables */
    double *c1, *A[3], *bup, *bdn, *ain, *aout, *Al[3];
    size_t *hi_at, *lo_at; /* index into the (c^n | c1) pair: idx + n*field */
} fwd_t;

static void fwd_free(fwd_t* f)
{
    free(f->xm); free(f->xp); free(f->ym); free(f->yp); free(f->zm); free(f->zp);
    free(f->c1); free(f->bup); free(f->bdn); free(f->ain); free(f->aout);
    for (int d = 0; d < 3; d++) { free(f->A[d]); free(f->Al[d]); }
    free(f->hi_at); free(f->lo_at);
}

static int* wraptab(int n, int off)
{
    int* t = malloc(sizeof(int) * n);
    for (int i = 0; t && i < n; i++) t[i] = ((i + off) % n + n) % n;
    return t;
}

/* Forward step with every intermediate kept (the limiter part only when nonosc). */
static int fwd(fwd_t* f, int nx, int ny, int nz, const double* U, const double* V,
               const double* W, int nonosc, const double* c0)
{
    memset(f, 0, sizeof *f);
    f->nx = nx; f->ny = ny; f->nz = nz;
    size_t n = f->n = (size_t)nx * ny * nz;
    f->xm = wraptab(nx, -1); f->xp = wraptab(nx, 1);
    f->ym = wraptab(ny, -1); f->yp = wraptab(ny, 1);
    f->zm = wraptab(nz, -1); f->zp = wraptab(nz, 1);
    f->c1 = malloc(8 * n); f->bup = malloc(8 * n); f->bdn = malloc(8 * n);
    f->ain = malloc(8 * n); f->aout = malloc(8 * n);
    f->hi_at = malloc(sizeof(size_t) * n); f->lo_at = malloc(sizeof(size_t) * n);
    for (int d = 0; d < 3; d++) { f->A[d] = malloc(8 * n); f->Al[d] = malloc(8 * n); }
    if (!f->xm || !f->xp || !f->ym || !f->yp || !f->zm || !f->zp || !f->c1 || !f->bup ||
        !f->bdn || !f->ain || !f->aout || !f->hi_at || !f->lo_at || !f->A[0] || !f->A[1] ||
        !f->A[2] || !f->Al[0] || !f->Al[1] || !f->Al[2])
        return 1;
    orc_per_donor(nx, ny, nz, U, V, W, c0, f->c1);
    orc_per_antidiff(nx, ny, nz, f->c1, U, V, W, f->A[0], f->A[1], f->A[2]);
    for (int d = 0; d < 3; d++) memcpy(f->Al[d], f->A[d], 8 * n);
    if (!nonosc) return 0;
    const double* cp = f->c1;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = PI(i, j, k);
                size_t xl = PI(f->xm[i], j, k), xr = PI(f->xp[i], j, k);
                size_t yd = PI(i, f->ym[j], k), yu = PI(i, f->yp[j], k);
                size_t zb = PI(i, j, f->zm[k]), zf = PI(i, j, f->zp[k]);
                size_t nb[7] = {o, xl, xr, yd, yu, zb, zf};
                /* the same selection sequence as orc_per_limiter_beta, tracking which
                 * element (cell, field) each running max/min holds */
                double hi = c0[o], lo = c0[o];
                size_t hat = o, lat = o;
                for (int q = 0; q < 7; q++) {
                    double a = c0[nb[q]], b = cp[nb[q]];
                    double mh = a > b ? a : b;
                    size_t mha = a > b ? nb[q] : nb[q] + n;
                    if (!(hi > mh)) { hi = mh; hat = mha; }
                    double ml = a < b ? a : b;
                    size_t mla = a < b ? nb[q] : nb[q] + n;
                    if (!(lo < ml)) { lo = ml; lat = mla; }
                }
                double c = cp[o];
                double fxl = 0.5 * F2(cp[xl], c, f->A[0][xl]), fxr = 0.5 * F2(c, cp[xr], f->A[0][o]);
                double fyl = 0.5 * F2(cp[yd], c, f->A[1][yd]), fyr = 0.5 * F2(c, cp[yu], f->A[1][o]);
                double fzl = 0.5 * F2(cp[zb], c, f->A[2][zb]), fzr = 0.5 * F2(c, cp[zf], f->A[2][o]);
#define POS(a) ((a) > 0.0 ? (a) : 0.0)
#define NEG(a) ((a) < 0.0 ? (a) : 0.0)
                double ain = ((POS(fxl) - NEG(fxr)) + (POS(fyl) - NEG(fyr))) + (POS(fzl) - NEG(fzr));
                double aout = ((POS(fxr) - NEG(fxl)) + (POS(fyr) - NEG(fyl))) + (POS(fzr) - NEG(fzl));
                f->ain[o] = ain;
                f->aout[o] = aout;
                f->bup[o] = (hi - c) / (ain + MINV);
                f->bdn[o] = (c - lo) / (aout + MINV);
                f->hi_at[o] = hat;
                f->lo_at[o] = lat;
            }
    orc_per_limiter_apply(nx, ny, nz, f->bup, f->bdn, f->Al[0], f->Al[1], f->Al[2]);
    return 0;
}

/* Which ratio the limiter selected for a face: 0 = the constant 1, 1 = the first
 * argument's ratio, 2 = the second's; mirrors mn2(1, mn2(a, b)). */
static inline int pick(double a, double b)
{
    double m = a < b ? a : b;
    int w = a < b ? 1 : 2;
    return 1.0 < m ? 0 : w;
}

/* x-face antidiffusive velocity of cell o (face i+1/2) as a function of c1 only (the
 * velocities of the pass are the fixed Courant numbers): the partial derivatives with
 * respect to its ten cells.  Order: p (=cr), m (=cc), then the up/down pairs of the two
 * transverse ratios.  Returns through g[10] with cells in idx[10]. */
static void antidiff_partials(const double* c, double a, double s1, double s2, size_t ip,
                              size_t im, const size_t up1[2], const size_t dn1[2],
                              const size_t up2[2], const size_t dn2[2], double g[10],
                              size_t idx[10])
{
    double p = c[ip], m = c[im];
    double K = ab(a) * (1.0 - ab(a));
    double den = p + m + MINV;
    /* (x/den)/den, not x/(den*den): den can be ~DBL_MIN where c vanishes, and its
     * square would underflow to zero */
    g[0] = K * (((2.0 * m + MINV) / den) / den);
    g[1] = -K * (((2.0 * p + MINV) / den) / den);
    idx[0] = ip;
    idx[1] = im;
    double n1 = c[up1[0]] + c[up1[1]] - c[dn1[0]] - c[dn1[1]];
    double D1 = c[up1[0]] + c[up1[1]] + c[dn1[0]] + c[dn1[1]] + MINV;
    double n2 = c[up2[0]] + c[up2[1]] - c[dn2[0]] - c[dn2[1]];
    double D2 = c[up2[0]] + c[up2[1]] + c[dn2[0]] + c[dn2[1]] + MINV;
    double gb = -0.5 * a; /* dA/dB = dA/dC */
    double gu1 = gb * s1 * (((D1 - n1) / D1) / D1), gd1 = gb * s1 * (((-D1 - n1) / D1) / D1);
    double gu2 = gb * s2 * (((D2 - n2) / D2) / D2), gd2 = gb * s2 * (((-D2 - n2) / D2) / D2);
    g[2] = g[3] = gu1; idx[2] = up1[0]; idx[3] = up1[1];
    g[4] = g[5] = gd1; idx[4] = dn1[0]; idx[5] = dn1[1];
    g[6] = g[7] = gu2; idx[6] = up2[0]; idx[7] = up2[1];
    g[8] = g[9] = gd2; idx[8] = dn2[0]; idx[9] = dn2[1];
}

/* The partials of all three face families of cell (i,j,k); fam 0/1/2 = x/y/z face.
 * Stencils and strokes as orc_per_antidiff (uu.c, vv.c, ww.c interior formulas). */
static void face_partials(const fwd_t* f, const double* U, const double* V,
                          const double* W, int i, int j, int k, int fam, double g[10],
                          size_t idx[10])
{
    const int nx = f->nx, ny = f->ny;
    const int im = f->xm[i], ip = f->xp[i], jm = f->ym[j], jp = f->yp[j];
    const int km = f->zm[k], kp = f->zp[k];
    const size_t o = PI(i, j, k);
    if (fam == 0) {
        double s1 = 0.25 * (V[PI(ip, jm, k)] + V[PI(ip, j, k)] + V[o] + V[PI(i, jm, k)]);
        double s2 = 0.25 * (W[PI(ip, j, km)] + W[PI(ip, j, k)] + W[o] + W[PI(i, j, km)]);
        size_t up1[2] = {PI(ip, jp, k), PI(i, jp, k)}, dn1[2] = {PI(i, jm, k), PI(ip, jm, k)};
        size_t up2[2] = {PI(ip, j, kp), PI(i, j, kp)}, dn2[2] = {PI(i, j, km), PI(ip, j, km)};
        antidiff_partials(f->c1, U[o], s1, s2, PI(ip, j, k), o, up1, dn1, up2, dn2, g, idx);
    } else if (fam == 1) {
        double s1 = 0.25 * (U[PI(im, jp, k)] + U[PI(i, jp, k)] + U[o] + U[PI(im, j, k)]);
        double s2 = 0.25 * (W[PI(i, jp, km)] + W[PI(i, jp, k)] + W[o] + W[PI(i, j, km)]);
        size_t up1[2] = {PI(ip, jp, k), PI(ip, j, k)}, dn1[2] = {PI(im, j, k), PI(im, jp, k)};
        size_t up2[2] = {PI(i, jp, kp), PI(i, j, kp)}, dn2[2] = {PI(i, j, km), PI(i, jp, km)};
        antidiff_partials(f->c1, V[o], s1, s2, PI(i, jp, k), o, up1, dn1, up2, dn2, g, idx);
    } else {
        double s1 = 0.25 * (U[o] + U[PI(i, j, kp)] + U[PI(im, j, kp)] + U[PI(im, j, k)]);
        double s2 = 0.25 * (V[PI(i, jm, kp)] + V[PI(i, j, kp)] + V[o] + V[PI(i, jm, k)]);
        size_t up1[2] = {PI(ip, j, kp), PI(ip, j, k)}, dn1[2] = {PI(im, j, k), PI(im, j, kp)};
        size_t up2[2] = {PI(i, jp, kp), PI(i, jp, k)}, dn2[2] = {PI(i, jm, k), PI(i, jm, kp)};
        antidiff_partials(f->c1, W[o], s1, s2, PI(i, j, kp), o, up1, dn1, up2, dn2, g, idx);
    }
}

/* Partials of f = 0.5*F2(l, r, v): d/dl, d/dr, d/dv. */
static inline void flux_partials(double l, double r, double v, double* gl, double* gr,
                                 double* gv)
{
    double s = sg(v);
    *gl = 0.5 * (v + ab(v));
    *gr = 0.5 * (v - ab(v));
    *gv = 0.5 * (l * (1.0 + s) + r * (1.0 - s));
}

int orc_tlm_step(int nx, int ny, int nz, const double* U, const double* V, const double* W,
                 int nonosc, const double* c0, const double* dc0, const double* dq,
                 double* dout)
{
    fwd_t f;
    if (fwd(&f, nx, ny, nz, U, V, W, nonosc, c0)) { fwd_free(&f); return 1; }
    size_t n = f.n;
    double* dc1 = malloc(8 * n);
    double* dA[3] = {calloc(n, 8), calloc(n, 8), calloc(n, 8)};
    double* dAl[3] = {malloc(8 * n), malloc(8 * n), malloc(8 * n)};
    double *dbu = calloc(n, 8), *dbd = calloc(n, 8);
    int rc = 1;
    if (!dc1 || !dA[0] || !dA[1] || !dA[2] || !dAl[0] || !dAl[1] || !dAl[2] || !dbu || !dbd)
        goto done;
    /* donor pass: linear in c */
    orc_per_donor(nx, ny, nz, U, V, W, dc0, dc1);
    /* antidiffusive velocities */
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                for (int fam = 0; fam < 3; fam++) {
                    double g[10];
                    size_t idx[10];
                    face_partials(&f, U, V, W, i, j, k, fam, g, idx);
                    double s = 0.0;
                    for (int q = 0; q < 10; q++) s += g[q] * dc1[idx[q]];
                    dA[fam][PI(i, j, k)] = s;
                }
    if (nonosc) {
        const double* cp = f.c1;
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    size_t o = PI(i, j, k);
                    size_t nbl[3] = {PI(f.xm[i], j, k), PI(i, f.ym[j], k), PI(i, j, f.zm[k])};
                    size_t nbr[3] = {PI(f.xp[i], j, k), PI(i, f.yp[j], k), PI(i, j, f.zp[k])};
                    double dain = 0.0, daout = 0.0;
                    for (int d = 0; d < 3; d++) {
                        double gl, gr, gv;
                        /* lower face: flux from nbl into o */
                        double fl = 0.5 * F2(cp[nbl[d]], cp[o], f.A[d][nbl[d]]);
                        flux_partials(cp[nbl[d]], cp[o], f.A[d][nbl[d]], &gl, &gr, &gv);
                        double dfl = gl * dc1[nbl[d]] + gr * dc1[o] + gv * dA[d][nbl[d]];
                        double fr = 0.5 * F2(cp[o], cp[nbr[d]], f.A[d][o]);
                        flux_partials(cp[o], cp[nbr[d]], f.A[d][o], &gl, &gr, &gv);
                        double dfr = gl * dc1[o] + gr * dc1[nbr[d]] + gv * dA[d][o];
                        if (fl > 0) dain += dfl;
                        if (fr < 0) dain -= dfr;
                        if (fr > 0) daout += dfr;
                        if (fl < 0) daout -= dfl;
                    }
                    double dhi = f.hi_at[o] < n ? dc0[f.hi_at[o]] : dc1[f.hi_at[o] - n];
                    double dlo = f.lo_at[o] < n ? dc0[f.lo_at[o]] : dc1[f.lo_at[o] - n];
                    double eu = f.hi_at[o] == o + n ? 0.0 : dhi - dc1[o];
                    double ed = f.lo_at[o] == o + n ? 0.0 : dc1[o] - dlo;
                    /* a ratio the limiter did not select may be inf: its derivative is
                     * never used, keep it finite */
                    dbu[o] = f.bup[o] <= 1.0 ? (eu - f.bup[o] * dain) / (f.ain[o] + MINV) : 0.0;
                    dbd[o] = f.bdn[o] <= 1.0 ? (ed - f.bdn[o] * daout) / (f.aout[o] + MINV) : 0.0;
                }
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    size_t o = PI(i, j, k);
                    size_t up[3] = {PI(f.xp[i], j, k), PI(i, f.yp[j], k), PI(i, j, f.zp[k])};
                    for (int d = 0; d < 3; d++) {
                        double a = f.A[d][o];
                        size_t r = up[d];
                        double da = dA[d][o], v;
                        if (a > 0) {
                            int w = pick(f.bdn[o], f.bup[r]);
                            double m = w == 0 ? 1.0 : (w == 1 ? f.bdn[o] : f.bup[r]);
                            double dm = w == 0 ? 0.0 : (w == 1 ? dbd[o] : dbu[r]);
                            v = m * da + a * dm;
                        } else if (a < 0) {
                            int w = pick(f.bup[o], f.bdn[r]);
                            double m = w == 0 ? 1.0 : (w == 1 ? f.bup[o] : f.bdn[r]);
                            double dm = w == 0 ? 0.0 : (w == 1 ? dbu[o] : dbd[r]);
                            v = m * da + a * dm;
                        } else {
                            v = 0.0;
                        }
                        dAl[d][o] = v;
                    }
                }
    } else {
        for (int d = 0; d < 3; d++) memcpy(dAl[d], dA[d], 8 * n);
    }
    /* corrective donor pass: bilinear in (c1, Al) */
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = PI(i, j, k);
                size_t nbl[3] = {PI(f.xm[i], j, k), PI(i, f.ym[j], k), PI(i, j, f.zm[k])};
                size_t nbr[3] = {PI(f.xp[i], j, k), PI(i, f.yp[j], k), PI(i, j, f.zp[k])};
                double s = 0.0;
                for (int d = 0; d < 3; d++) {
                    double gl, gr, gv;
                    flux_partials(f.c1[o], f.c1[nbr[d]], f.Al[d][o], &gl, &gr, &gv);
                    double dfr = gl * dc1[o] + gr * dc1[nbr[d]] + gv * dAl[d][o];
                    flux_partials(f.c1[nbl[d]], f.c1[o], f.Al[d][nbl[d]], &gl, &gr, &gv);
                    double dfl = gl * dc1[nbl[d]] + gr * dc1[o] + gv * dAl[d][nbl[d]];
                    s += dfr - dfl;
                }
                /* out = c1 - 0.5*sum(F2 differences) = c1 - sum(flux differences) */
                dout[o] = dc1[o] - s + (dq ? dq[o] : 0.0);
            }
    rc = 0;
done:
    free(dc1); free(dbu); free(dbd);
    for (int d = 0; d < 3; d++) { free(dA[d]); free(dAl[d]); }
    fwd_free(&f);
    return rc;
}

int orc_adj_step(int nx, int ny, int nz, const double* U, const double* V, const double* W,
                 int nonosc, const double* c0, const double* lout, double* lin,
                 double* lq)
{
    fwd_t f;
    if (fwd(&f, nx, ny, nz, U, V, W, nonosc, c0)) { fwd_free(&f); return 1; }
    size_t n = f.n;
    double* l1 = calloc(n, 8);
    double* lA[3] = {calloc(n, 8), calloc(n, 8), calloc(n, 8)};
    double* lAl[3] = {calloc(n, 8), calloc(n, 8), calloc(n, 8)};
    double *lbu = calloc(n, 8), *lbd = calloc(n, 8);
    int rc = 1;
    if (!l1 || !lA[0] || !lA[1] || !lA[2] || !lAl[0] || !lAl[1] || !lAl[2] || !lbu || !lbd)
        goto done;
    if (lq)
        for (size_t o = 0; o < n; o++) lq[o] += lout[o];
    for (size_t o = 0; o < n; o++) lin[o] = 0.0;
    /* corrective donor pass, transposed */
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = PI(i, j, k);
                size_t nbl[3] = {PI(f.xm[i], j, k), PI(i, f.ym[j], k), PI(i, j, f.zm[k])};
                size_t nbr[3] = {PI(f.xp[i], j, k), PI(i, f.yp[j], k), PI(i, j, f.zp[k])};
                double L = lout[o];
                l1[o] += L;
                for (int d = 0; d < 3; d++) {
                    double gl, gr, gv;
                    flux_partials(f.c1[o], f.c1[nbr[d]], f.Al[d][o], &gl, &gr, &gv);
                    l1[o] -= gl * L; l1[nbr[d]] -= gr * L; lAl[d][o] -= gv * L;
                    flux_partials(f.c1[nbl[d]], f.c1[o], f.Al[d][nbl[d]], &gl, &gr, &gv);
                    l1[nbl[d]] += gl * L; l1[o] += gr * L; lAl[d][nbl[d]] += gv * L;
                }
            }
    if (nonosc) {
        /* limited faces -> unlimited faces and ratios */
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    size_t o = PI(i, j, k);
                    size_t up[3] = {PI(f.xp[i], j, k), PI(i, f.yp[j], k), PI(i, j, f.zp[k])};
                    for (int d = 0; d < 3; d++) {
                        double a = f.A[d][o], L = lAl[d][o];
                        size_t r = up[d];
                        if (a > 0) {
                            int w = pick(f.bdn[o], f.bup[r]);
                            double m = w == 0 ? 1.0 : (w == 1 ? f.bdn[o] : f.bup[r]);
                            lA[d][o] += m * L;
                            if (w == 1) lbd[o] += a * L;
                            if (w == 2) lbu[r] += a * L;
                        } else if (a < 0) {
                            int w = pick(f.bup[o], f.bdn[r]);
                            double m = w == 0 ? 1.0 : (w == 1 ? f.bup[o] : f.bdn[r]);
                            lA[d][o] += m * L;
                            if (w == 1) lbu[o] += a * L;
                            if (w == 2) lbd[r] += a * L;
                        }
                    }
                }
        /* ratios -> bounds, fluxes, c1, c^n */
        const double* cp = f.c1;
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    size_t o = PI(i, j, k);
                    size_t nbl[3] = {PI(f.xm[i], j, k), PI(i, f.ym[j], k), PI(i, j, f.zm[k])};
                    size_t nbr[3] = {PI(f.xp[i], j, k), PI(i, f.yp[j], k), PI(i, j, f.zp[k])};
                    /* bup = (hi - c)/(ain+e), bdn = (c - lo)/(aout+e).  A ratio that was
                     * not selected carries no adjoint (and may be inf).  When the bound is
                     * the cell's own c1 value the two terms are one and the same and cancel
                     * exactly -- folded here, since each alone can be ~1/DBL_MIN. */
                    double lain = 0.0, laout = 0.0;
                    if (lbu[o] != 0.0) {
                        double iu = lbu[o] / (f.ain[o] + MINV);
                        size_t h = f.hi_at[o];
                        if (h != o + n) {
                            if (h < n) lin[h] += iu; else l1[h - n] += iu;
                            l1[o] -= iu;
                        }
                        lain = -iu * f.bup[o];
                    }
                    if (lbd[o] != 0.0) {
                        double id = lbd[o] / (f.aout[o] + MINV);
                        size_t lo = f.lo_at[o];
                        if (lo != o + n) {
                            if (lo < n) lin[lo] -= id; else l1[lo - n] -= id;
                            l1[o] += id;
                        }
                        laout = -id * f.bdn[o];
                    }
                    for (int d = 0; d < 3; d++) {
                        double gl, gr, gv;
                        double fl = 0.5 * F2(cp[nbl[d]], cp[o], f.A[d][nbl[d]]);
                        double fr = 0.5 * F2(cp[o], cp[nbr[d]], f.A[d][o]);
                        double lfl = (fl > 0 ? lain : 0.0) - (fl < 0 ? laout : 0.0);
                        double lfr = (fr > 0 ? laout : 0.0) - (fr < 0 ? lain : 0.0);
                        flux_partials(cp[nbl[d]], cp[o], f.A[d][nbl[d]], &gl, &gr, &gv);
                        l1[nbl[d]] += gl * lfl; l1[o] += gr * lfl; lA[d][nbl[d]] += gv * lfl;
                        flux_partials(cp[o], cp[nbr[d]], f.A[d][o], &gl, &gr, &gv);
                        l1[o] += gl * lfr; l1[nbr[d]] += gr * lfr; lA[d][o] += gv * lfr;
                    }
                }
    } else {
        for (int d = 0; d < 3; d++) memcpy(lA[d], lAl[d], 8 * n);
    }
    /* antidiffusive velocities -> c1 */
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                for (int fam = 0; fam < 3; fam++) {
                    double g[10];
                    size_t idx[10];
                    face_partials(&f, U, V, W, i, j, k, fam, g, idx);
                    double L = lA[fam][PI(i, j, k)];
                    for (int q = 0; q < 10; q++) l1[idx[q]] += g[q] * L;
                }
    /* donor pass, transposed: c1 = c0 - sum over faces of 0.5*F2 differences */
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = PI(i, j, k);
                size_t nbl[3] = {PI(f.xm[i], j, k), PI(i, f.ym[j], k), PI(i, j, f.zm[k])};
                size_t nbr[3] = {PI(f.xp[i], j, k), PI(i, f.yp[j], k), PI(i, j, f.zp[k])};
                const double* VV[3] = {U, V, W};
                double L = l1[o];
                lin[o] += L;
                for (int d = 0; d < 3; d++) {
                    double gl, gr, gv;
                    flux_partials(0.0, 0.0, VV[d][o], &gl, &gr, &gv);
                    lin[o] -= gl * L; lin[nbr[d]] -= gr * L;
                    flux_partials(0.0, 0.0, VV[d][nbl[d]], &gl, &gr, &gv);
                    lin[nbl[d]] += gl * L; lin[o] += gr * L;
                }
            }
    rc = 0;
done:
    free(l1); free(lbu); free(lbd);
    for (int d = 0; d < 3; d++) { free(lA[d]); free(lAl[d]); }
    fwd_free(&f);
    return rc;
}
