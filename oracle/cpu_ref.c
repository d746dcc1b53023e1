/* CPU oracle for the MPDATA transport step -- TEST INFRASTRUCTURE ONLY (see cpu_ref.h).
 *
 * FRAME functions restate the reference's core/ arithmetic in its own evaluation order;
 * each cites the reference lines it follows.  Compile with -ffp-contract=off.
 */
#include "cpu_ref.h"

#include <float.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <string.h>

/* smolar_int.h:25 -- guard added to every ratio denominator. */
#define MINV DBL_MIN

/* smolar_int.h:8 -- abs(a) = (a > 0 ? a : -a).  Differs from fabs only in the sign of
 * a zero result, which never changes a non-zero value downstream. */
static inline double ab(double a) { return a > 0 ? a : -a; }

/* smolar_pass.c:104 -- twice the donor-cell flux through a face with Courant number v
 * between left/lower cell value l and right/upper cell value r. */
static inline double F2(double l, double r, double v)
{
    double av = ab(v);
    return l * (v + av) + r * (v - av);
}

/* ============================ FRAME mode ========================================= */

#define CI(i, j, k) ((size_t)(k) * nx * ny + (size_t)(j) * nx + (i))
#define UI(i, j, k) ((size_t)((k) - 1) * (nx - 1) * (ny - 2) + (size_t)((j) - 1) * (nx - 1) + (i))
#define VI(i, j, k) ((size_t)((k) - 1) * (nx - 2) * (ny - 1) + (size_t)(j) * (nx - 2) + (i) - 1)
#define WI(i, j, k) ((size_t)(k) * (nx - 2) * (ny - 2) + (size_t)((j) - 1) * (nx - 2) + (i) - 1)

/* uu.c:126-138, vv.c:137-149, ww.c:303-315: face Courant numbers,
 * (0.25*dt/dx[i]) * (((a_i + a_{i+1}) + b_i) + b_{i+1}), spacing of the lower cell. */
void orc_frame_courant(int nx, int ny, int ns,
                       const double* dx, const double* dy, const double* ds, double dt,
                       const double* u0, const double* u1, const double* v0,
                       const double* v1, const double* w0, const double* w1,
                       double* U, double* V, double* W)
{
    for (int k = 1; k < ns - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 0; i < nx - 1; i++) {
                double q = 0.25 * dt / dx[i];
                U[UI(i, j, k)] = q * (((u0[CI(i, j, k)] + u0[CI(i + 1, j, k)]) +
                                       u1[CI(i, j, k)]) + u1[CI(i + 1, j, k)]);
            }
    for (int k = 1; k < ns - 1; k++)
        for (int j = 0; j < ny - 1; j++) {
            double q = 0.25 * dt / dy[j];
            for (int i = 1; i < nx - 1; i++)
                V[VI(i, j, k)] = q * (((v0[CI(i, j, k)] + v0[CI(i, j + 1, k)]) +
                                       v1[CI(i, j, k)]) + v1[CI(i, j + 1, k)]);
        }
    for (int k = 0; k < ns - 1; k++) {
        double q = 0.25 * dt / ds[k];
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++)
                W[WI(i, j, k)] = q * (((w0[CI(i, j, k)] + w0[CI(i, j, k + 1)]) +
                                       w1[CI(i, j, k)]) + w1[CI(i, j, k + 1)]);
    }
}

/* smolar.c:5-31 -- copy the outer one-cell layer of every face. */
void orc_frame_bndcpy(int nx, int ny, int ns, double* to, const double* from)
{
    for (int k = 0; k < ns; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                if (k == 0 || k == ns - 1 || j == 0 || j == ny - 1 || i == 0 || i == nx - 1)
                    to[CI(i, j, k)] = from[CI(i, j, k)];
}

/* smolar_pass.c:30-58 -- donor-cell update of the interior. */
void orc_frame_donor(int nx, int ny, int ns, const double* U, const double* V,
                     const double* W, const double* c0, double* c1)
{
    for (int k = 1; k < ns - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                double c = c0[CI(i, j, k)];
                double fx = F2(c, c0[CI(i + 1, j, k)], U[UI(i, j, k)]) -
                            F2(c0[CI(i - 1, j, k)], c, U[UI(i - 1, j, k)]);
                double fy = F2(c, c0[CI(i, j + 1, k)], V[VI(i, j, k)]) -
                            F2(c0[CI(i, j - 1, k)], c, V[VI(i, j - 1, k)]);
                double fz = F2(c, c0[CI(i, j, k + 1)], W[WI(i, j, k)]) -
                            F2(c0[CI(i, j, k - 1)], c, W[WI(i, j, k - 1)]);
                c1[CI(i, j, k)] = c - 0.5 * ((fx + fy) + fz);
            }
}

/* The antidiffusive pseudo-velocity of one face (uu.c:18-32, vv.c:21-35, ww.c:180-194):
 *   ((-div + |a|(1-|a|)(p-m)/(p+m+MINV)) - 0.5*(a*(s1*n1/d1 + s2*n2/d2)))
 * a = face Courant number of the previous pass, p/m = cell values on either side,
 * s1/s2 = averaged transverse velocities, n/d = the transverse gradient ratios. */
static inline double antidiff(double div, double a, double p, double m,
                              double s1, double n1, double d1, double s2, double n2,
                              double d2)
{
    double r = (ab(a) * (1.0 - ab(a))) * (p - m) / (p + m + MINV);
    return (-div + r) - 0.5 * (a * (s1 * n1 / (d1 + MINV) + (s2 * n2 / (d2 + MINV))));
}

/* uu.c:56-99 -- x-face antidiffusive velocities; reduced patterns (div=0, two-face
 * transverse averages of the interior neighbour) on the faces next to the frame. */
void orc_frame_antidiff_u(int nx, int ny, int ns, const double* c, const double* U0,
                          const double* V0, const double* W0, double* U1)
{
    for (int k = 1; k < ns - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 0; i < nx - 1; i++) {
                double a = U0[UI(i, j, k)], div, sv, sw;
                if (i == 0 || i == nx - 2) {
                    int ii = (i == 0) ? 1 : nx - 2;
                    div = 0.0;
                    sv = 0.5 * (V0[VI(ii, j - 1, k)] + V0[VI(ii, j, k)]);
                    sw = 0.5 * (W0[WI(ii, j, k - 1)] + W0[WI(ii, j, k)]);
                } else {
                    double vr = V0[VI(i + 1, j - 1, k)], vru = V0[VI(i + 1, j, k)];
                    double vu = V0[VI(i, j, k)], vc = V0[VI(i, j - 1, k)];
                    double wr = W0[WI(i + 1, j, k - 1)], wrf = W0[WI(i + 1, j, k)];
                    double wf = W0[WI(i, j, k)], wc = W0[WI(i, j, k - 1)];
                    sv = 0.25 * (vr + vru + vu + vc);
                    sw = 0.25 * (wr + wrf + wf + wc);
                    div = 0.25 * a * (U0[UI(i + 1, j, k)] - U0[UI(i - 1, j, k)] +
                                      vr - vru - vu + vc + wr - wrf - wf + wc);
                }
                double cc = c[CI(i, j, k)], cr = c[CI(i + 1, j, k)];
                double cru = c[CI(i + 1, j + 1, k)], cu = c[CI(i, j + 1, k)];
                double cd = c[CI(i, j - 1, k)], crd = c[CI(i + 1, j - 1, k)];
                double crf = c[CI(i + 1, j, k + 1)], cf = c[CI(i, j, k + 1)];
                double cb = c[CI(i, j, k - 1)], crb = c[CI(i + 1, j, k - 1)];
                U1[UI(i, j, k)] = antidiff(div, a, cr, cc,
                                           sv, cru + cu - cd - crd, cru + cu + cd + crd,
                                           sw, crf + cf - cb - crb, crf + cf + cb + crb);
            }
}

/* vv.c:58-104 -- y-face antidiffusive velocities.  The interior loop starts at i=2
 * (vv.c:81-82), so column i=1 of faces j=1..ny-3 is left as the buffer held it.  The
 * divergence term takes its x-velocity difference one cell to the left (vv.c:87-89). */
void orc_frame_antidiff_v(int nx, int ny, int ns, const double* c, const double* U0,
                          const double* V0, const double* W0, double* V1)
{
    for (int k = 1; k < ns - 1; k++)
        for (int j = 0; j < ny - 1; j++) {
            int edge = (j == 0 || j == ny - 2);
            for (int i = edge ? 1 : 2; i < nx - 1; i++) {
                double a = V0[VI(i, j, k)], div, su, sw;
                if (edge) {
                    int jj = (j == 0) ? 1 : ny - 2;
                    div = 0.0;
                    su = 0.5 * (U0[UI(i - 1, jj, k)] + U0[UI(i, jj, k)]);
                    sw = 0.5 * (W0[WI(i, jj, k - 1)] + W0[WI(i, jj, k)]);
                } else {
                    double uu = U0[UI(i - 1, j + 1, k)], uru = U0[UI(i, j + 1, k)];
                    double ur = U0[UI(i, j, k)], uc = U0[UI(i - 1, j, k)];
                    double ulu = U0[UI(i - 2, j + 1, k)], ul = U0[UI(i - 2, j, k)];
                    double wu = W0[WI(i, j + 1, k - 1)], wuf = W0[WI(i, j + 1, k)];
                    double wf = W0[WI(i, j, k)], wc = W0[WI(i, j, k - 1)];
                    su = 0.25 * (uu + uru + ur + uc);
                    sw = 0.25 * (wu + wuf + wf + wc);
                    div = 0.25 * a * (V0[VI(i, j + 1, k)] - V0[VI(i, j - 1, k)] +
                                      wu - wuf - wf + wc + uu - ulu - ul + uc);
                }
                double cc = c[CI(i, j, k)], cu = c[CI(i, j + 1, k)];
                double cru = c[CI(i + 1, j + 1, k)], cr = c[CI(i + 1, j, k)];
                double cl = c[CI(i - 1, j, k)], clu = c[CI(i - 1, j + 1, k)];
                double cuf = c[CI(i, j + 1, k + 1)], cf = c[CI(i, j, k + 1)];
                double cb = c[CI(i, j, k - 1)], cub = c[CI(i, j + 1, k - 1)];
                V1[VI(i, j, k)] = antidiff(div, a, cu, cc,
                                           su, cru + cr - cl - clu, cru + cr + cl + clu,
                                           sw, cuf + cf - cb - cub, cuf + cf + cb + cub);
            }
        }
}

/* ww.c:222-270 -- z-face antidiffusive velocities. */
void orc_frame_antidiff_w(int nx, int ny, int ns, const double* c, const double* U0,
                          const double* V0, const double* W0, double* W1)
{
    for (int k = 0; k < ns - 1; k++)
        for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
                double a = W0[WI(i, j, k)], div, su, sv;
                if (k == 0 || k == ns - 2) {
                    int kk = (k == 0) ? 1 : ns - 2;
                    div = 0.0;
                    su = 0.5 * (U0[UI(i - 1, j, kk)] + U0[UI(i, j, kk)]);
                    sv = 0.5 * (V0[VI(i, j - 1, kk)] + V0[VI(i, j, kk)]);
                } else {
                    double ur = U0[UI(i, j, k)], urf = U0[UI(i, j, k + 1)];
                    double uf = U0[UI(i - 1, j, k + 1)], uc = U0[UI(i - 1, j, k)];
                    double vf = V0[VI(i, j - 1, k + 1)], vuf = V0[VI(i, j, k + 1)];
                    double vu = V0[VI(i, j, k)], vc = V0[VI(i, j - 1, k)];
                    su = 0.25 * (ur + urf + uf + uc);
                    sv = 0.25 * (vf + vuf + vu + vc);
                    div = 0.25 * a * (W0[WI(i, j, k + 1)] - W0[WI(i, j, k - 1)] +
                                      uf - urf - ur + uc + vf - vuf - vu + vc);
                }
                double cc = c[CI(i, j, k)], cf = c[CI(i, j, k + 1)];
                double crf = c[CI(i + 1, j, k + 1)], cr = c[CI(i + 1, j, k)];
                double cl = c[CI(i - 1, j, k)], clf = c[CI(i - 1, j, k + 1)];
                double cuf = c[CI(i, j + 1, k + 1)], cu = c[CI(i, j + 1, k)];
                double cd = c[CI(i, j - 1, k)], cdf = c[CI(i, j - 1, k + 1)];
                W1[WI(i, j, k)] = antidiff(div, a, cf, cc,
                                           su, crf + cr - cl - clf, crf + cr + cl + clf,
                                           sv, cuf + cu - cd - cdf, cuf + cu + cd + cdf);
            }
}

/* smolar.c:33-63 with pvelocity.c:3-45 -- one step of `order` passes.  Pseudo-velocities
 * ping-pong between two buffers per direction (uu.c:43-44); the second buffer starts
 * zero-filled (the zero-malloc contract for the column vv_pass never writes).  Every
 * corrective pass uses the previous pass's U,V,W for all three directions (Jacobi,
 * pvelocity.c:24-26).  On return *c1 holds the result and *c0 the previous pass's field
 * (smolar.c:58). */
int orc_smolar(int nx, int ny, int ns, double* dx, double* dy, double* ds, const double dt,
               double* u0, double* u1, double* v0, double* v1, double* w0, double* w1,
               double** c0, double** c1, int flow_type, int order)
{
    (void)flow_type; /* never read by the reference (uu.c:111) */
    size_t nu = (size_t)(nx - 1) * (ny - 2) * (ns - 2);
    size_t nv = (size_t)(nx - 2) * (ny - 1) * (ns - 2);
    size_t nw = (size_t)(nx - 2) * (ny - 2) * (ns - 1);
    double *ua = calloc(nu ? nu : 1, 8), *ub = calloc(nu ? nu : 1, 8);
    double *va = calloc(nv ? nv : 1, 8), *vb = calloc(nv ? nv : 1, 8);
    double *wa = calloc(nw ? nw : 1, 8), *wb = calloc(nw ? nw : 1, 8);
    if (!ua || !ub || !va || !vb || !wa || !wb) {
        free(ua); free(ub); free(va); free(vb); free(wa); free(wb);
        return 1;
    }
    orc_frame_courant(nx, ny, ns, dx, dy, ds, dt, u0, u1, v0, v1, w0, w1, ua, va, wa);

    double *cold = *c0, *cnew = *c1, *t;
    for (int o = 0; o < order; o++) {
        if (o) {
            orc_frame_antidiff_u(nx, ny, ns, cold, ua, va, wa, ub);
            orc_frame_antidiff_v(nx, ny, ns, cold, ua, va, wa, vb);
            orc_frame_antidiff_w(nx, ny, ns, cold, ua, va, wa, wb);
            t = ua; ua = ub; ub = t;
            t = va; va = vb; vb = t;
            t = wa; wa = wb; wb = t;
        }
        orc_frame_bndcpy(nx, ny, ns, cnew, cold);
        orc_frame_donor(nx, ny, ns, ua, va, wa, cold, cnew);
        t = cnew; cnew = cold; cold = t;
    }
    *c1 = cold;
    *c0 = cnew;
    free(ua); free(ub); free(va); free(vb); free(wa); free(wb);
    return 0;
}

#undef CI
#undef UI
#undef VI
#undef WI

/* ============================ PER mode =========================================== */

/* Wrap tables: n[0..n) -> index of the neighbour at offset -2, -1, +1 (periodic). */
typedef struct { int *m2, *m1, *p1; } wrap_t;

static int wrap_make(wrap_t* w, int n)
{
    w->m2 = malloc(sizeof(int) * n);
    w->m1 = malloc(sizeof(int) * n);
    w->p1 = malloc(sizeof(int) * n);
    if (!w->m2 || !w->m1 || !w->p1) return 1;
    for (int i = 0; i < n; i++) {
        w->m2[i] = ((i - 2) % n + n) % n;
        w->m1[i] = ((i - 1) % n + n) % n;
        w->p1[i] = (i + 1) % n;
    }
    return 0;
}

static void wrap_free(wrap_t* w)
{
    free(w->m2); free(w->m1); free(w->p1);
}

#define PI(i, j, k) ((size_t)(k) * nx * ny + (size_t)(j) * nx + (i))

/* Threads of the periodic loops below (OpenMP over z planes; every cell is computed by
 * the same expression whatever the thread count, so results do not depend on it). */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Face Courant numbers on the periodic grid, same formula and order as uu.c:132-134
 * (x-face i+1/2 uses dx[i]; the face after the last cell wraps to cell 0). */
void orc_per_courant(int nx, int ny, int nz,
                     const double* dx, const double* dy, const double* dz, double dt,
                     const double* u0, const double* u1, const double* v0,
                     const double* v1, const double* w0, const double* w1,
                     double* U, double* V, double* W)
{
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++) {
        int kp = (k + 1) % nz;
        for (int j = 0; j < ny; j++) {
            int jp = (j + 1) % ny;
            for (int i = 0; i < nx; i++) {
                int ip = (i + 1) % nx;
                size_t c = PI(i, j, k);
                double qx = 0.25 * dt / dx[i], qy = 0.25 * dt / dy[j], qz = 0.25 * dt / dz[k];
                U[c] = qx * (((u0[c] + u0[PI(ip, j, k)]) + u1[c]) + u1[PI(ip, j, k)]);
                V[c] = qy * (((v0[c] + v0[PI(i, jp, k)]) + v1[c]) + v1[PI(i, jp, k)]);
                W[c] = qz * (((w0[c] + w0[PI(i, j, kp)]) + w1[c]) + w1[PI(i, j, kp)]);
            }
        }
    }
}

/* Donor-cell update (smolar_pass.c:40-54) on the periodic grid. */
void orc_per_donor(int nx, int ny, int nz, const double* U, const double* V,
                   const double* W, const double* cin, double* cout)
{
    wrap_t X, Y, Z;
    if (wrap_make(&X, nx) || wrap_make(&Y, ny) || wrap_make(&Z, nz)) return;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = PI(i, j, k);
                size_t xl = PI(X.m1[i], j, k), xr = PI(X.p1[i], j, k);
                size_t yd = PI(i, Y.m1[j], k), yu = PI(i, Y.p1[j], k);
                size_t zb = PI(i, j, Z.m1[k]), zf = PI(i, j, Z.p1[k]);
                double c = cin[o];
                double fx = F2(c, cin[xr], U[o]) - F2(cin[xl], c, U[xl]);
                double fy = F2(c, cin[yu], V[o]) - F2(cin[yd], c, V[yd]);
                double fz = F2(c, cin[zf], W[o]) - F2(cin[zb], c, W[zb]);
                cout[o] = c - 0.5 * ((fx + fy) + fz);
            }
    wrap_free(&X); wrap_free(&Y); wrap_free(&Z);
}

/* Antidiffusive velocities of all three face families on the periodic grid: the
 * reference's interior formulas (uu.c:81-87, vv.c:85-92, ww.c:248-255) with wrapped
 * indices, including the vv divergence term's one-cell x shift. */
void orc_per_antidiff(int nx, int ny, int nz, const double* c, const double* U0,
                      const double* V0, const double* W0, double* U1, double* V1,
                      double* W1)
{
    wrap_t X, Y, Z;
    if (wrap_make(&X, nx) || wrap_make(&Y, ny) || wrap_make(&Z, nz)) return;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++) {
        int km = Z.m1[k], kp = Z.p1[k];
        for (int j = 0; j < ny; j++) {
            int jm = Y.m1[j], jp = Y.p1[j];
            for (int i = 0; i < nx; i++) {
                int im = X.m1[i], ip = X.p1[i], im2 = X.m2[i];
                size_t o = PI(i, j, k);
                /* x-face i+1/2 */
                {
                    double a = U0[o];
                    double vr = V0[PI(ip, jm, k)], vru = V0[PI(ip, j, k)];
                    double vu = V0[o], vc = V0[PI(i, jm, k)];
                    double wr = W0[PI(ip, j, km)], wrf = W0[PI(ip, j, k)];
                    double wf = W0[o], wc = W0[PI(i, j, km)];
                    double sv = 0.25 * (vr + vru + vu + vc);
                    double sw = 0.25 * (wr + wrf + wf + wc);
                    double div = 0.25 * a * (U0[PI(ip, j, k)] - U0[PI(im, j, k)] +
                                             vr - vru - vu + vc + wr - wrf - wf + wc);
                    double cc = c[o], cr = c[PI(ip, j, k)];
                    double cru = c[PI(ip, jp, k)], cu = c[PI(i, jp, k)];
                    double cd = c[PI(i, jm, k)], crd = c[PI(ip, jm, k)];
                    double crf = c[PI(ip, j, kp)], cf = c[PI(i, j, kp)];
                    double cb = c[PI(i, j, km)], crb = c[PI(ip, j, km)];
                    U1[o] = antidiff(div, a, cr, cc,
                                     sv, cru + cu - cd - crd, cru + cu + cd + crd,
                                     sw, crf + cf - cb - crb, crf + cf + cb + crb);
                }
                /* y-face j+1/2 */
                {
                    double a = V0[o];
                    double uu = U0[PI(im, jp, k)], uru = U0[PI(i, jp, k)];
                    double ur = U0[o], uc = U0[PI(im, j, k)];
                    double ulu = U0[PI(im2, jp, k)], ul = U0[PI(im2, j, k)];
                    double wu = W0[PI(i, jp, km)], wuf = W0[PI(i, jp, k)];
                    double wf = W0[o], wc = W0[PI(i, j, km)];
                    double su = 0.25 * (uu + uru + ur + uc);
                    double sw = 0.25 * (wu + wuf + wf + wc);
                    double div = 0.25 * a * (V0[PI(i, jp, k)] - V0[PI(i, jm, k)] +
                                             wu - wuf - wf + wc + uu - ulu - ul + uc);
                    double cc = c[o], cu = c[PI(i, jp, k)];
                    double cru = c[PI(ip, jp, k)], cr = c[PI(ip, j, k)];
                    double cl = c[PI(im, j, k)], clu = c[PI(im, jp, k)];
                    double cuf = c[PI(i, jp, kp)], cf = c[PI(i, j, kp)];
                    double cb = c[PI(i, j, km)], cub = c[PI(i, jp, km)];
                    V1[o] = antidiff(div, a, cu, cc,
                                     su, cru + cr - cl - clu, cru + cr + cl + clu,
                                     sw, cuf + cf - cb - cub, cuf + cf + cb + cub);
                }
                /* z-face k+1/2 */
                {
                    double a = W0[o];
                    double ur = U0[o], urf = U0[PI(i, j, kp)];
                    double uf = U0[PI(im, j, kp)], uc = U0[PI(im, j, k)];
                    double vf = V0[PI(i, jm, kp)], vuf = V0[PI(i, j, kp)];
                    double vu = V0[o], vc = V0[PI(i, jm, k)];
                    double su = 0.25 * (ur + urf + uf + uc);
                    double sv = 0.25 * (vf + vuf + vu + vc);
                    double div = 0.25 * a * (W0[PI(i, j, kp)] - W0[PI(i, j, km)] +
                                             uf - urf - ur + uc + vf - vuf - vu + vc);
                    double cc = c[o], cf = c[PI(i, j, kp)];
                    double crf = c[PI(ip, j, kp)], cr = c[PI(ip, j, k)];
                    double cl = c[PI(im, j, k)], clf = c[PI(im, j, kp)];
                    double cuf = c[PI(i, jp, kp)], cu = c[PI(i, jp, k)];
                    double cd = c[PI(i, jm, k)], cdf = c[PI(i, jm, kp)];
                    W1[o] = antidiff(div, a, cf, cc,
                                     su, crf + cr - cl - clf, crf + cr + cl + clf,
                                     sv, cuf + cu - cd - cdf, cuf + cu + cd + cdf);
                }
            }
        }
    }
    wrap_free(&X); wrap_free(&Y); wrap_free(&Z);
}

/* ---- Nonoscillatory (flux-corrected) limiter --------------------------------------
 * Smolarkiewicz & Grabowski (1990), J. Comput. Phys. 86:355-375; in the form of
 * Smolarkiewicz & Margolin (1998), J. Comput. Phys. 140:459-480, eqs. (29)-(31).
 * Not present in the reference.  For a corrective pass acting on field cp with
 * antidiffusive face velocities U:
 *   cmax_i = max of c^n and cp over cell i and its six face neighbours (cmin likewise)
 *   f = 0.5*F2(c_left, c_right, U)   the pass's donor flux through a face
 *   Ain_i  = sum over axes of  pos(f_lower) - neg(f_upper)
 *   Aout_i = sum over axes of  pos(f_upper) - neg(f_lower)
 *   bup_i = (cmax_i - cp_i) / (Ain_i + MINV),   bdn_i = (cp_i - cmin_i) / (Aout_i + MINV)
 *   U_lim = min(1, bdn_i, bup_{i+1}) pos(U) + min(1, bup_i, bdn_{i+1}) neg(U)
 * MINV (= DBL_MIN, the reference's own ratio guard) keeps every ratio finite. */
static inline double pos(double a) { return a > 0.0 ? a : 0.0; }
static inline double neg(double a) { return a < 0.0 ? a : 0.0; }
static inline double mx2(double a, double b) { return a > b ? a : b; }
static inline double mn2(double a, double b) { return a < b ? a : b; }

void orc_per_limiter_beta(int nx, int ny, int nz, const double* cn, const double* cp,
                          const double* U, const double* V, const double* W,
                          double* bup, double* bdn)
{
    wrap_t X, Y, Z;
    if (wrap_make(&X, nx) || wrap_make(&Y, ny) || wrap_make(&Z, nz)) return;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = PI(i, j, k);
                size_t xl = PI(X.m1[i], j, k), xr = PI(X.p1[i], j, k);
                size_t yd = PI(i, Y.m1[j], k), yu = PI(i, Y.p1[j], k);
                size_t zb = PI(i, j, Z.m1[k]), zf = PI(i, j, Z.p1[k]);
                size_t nb[7] = {o, xl, xr, yd, yu, zb, zf};
                double hi = cn[o], lo = cn[o];
                for (int q = 0; q < 7; q++) {
                    hi = mx2(hi, mx2(cn[nb[q]], cp[nb[q]]));
                    lo = mn2(lo, mn2(cn[nb[q]], cp[nb[q]]));
                }
                double c = cp[o];
                double fxl = 0.5 * F2(cp[xl], c, U[xl]), fxr = 0.5 * F2(c, cp[xr], U[o]);
                double fyl = 0.5 * F2(cp[yd], c, V[yd]), fyr = 0.5 * F2(c, cp[yu], V[o]);
                double fzl = 0.5 * F2(cp[zb], c, W[zb]), fzr = 0.5 * F2(c, cp[zf], W[o]);
                double ain = ((pos(fxl) - neg(fxr)) + (pos(fyl) - neg(fyr))) +
                             (pos(fzl) - neg(fzr));
                double aout = ((pos(fxr) - neg(fxl)) + (pos(fyr) - neg(fyl))) +
                              (pos(fzr) - neg(fzl));
                bup[o] = (hi - c) / (ain + MINV);
                bdn[o] = (c - lo) / (aout + MINV);
            }
    wrap_free(&X); wrap_free(&Y); wrap_free(&Z);
}

static inline double limit_face(double u, double bup_l, double bdn_l, double bup_r,
                                double bdn_r)
{
    return mn2(1.0, mn2(bdn_l, bup_r)) * pos(u) + mn2(1.0, mn2(bup_l, bdn_r)) * neg(u);
}

void orc_per_limiter_apply(int nx, int ny, int nz, const double* bup, const double* bdn,
                           double* U, double* V, double* W)
{
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t o = PI(i, j, k);
                size_t xr = PI((i + 1) % nx, j, k), yu = PI(i, (j + 1) % ny, k);
                size_t zf = PI(i, j, (k + 1) % nz);
                U[o] = limit_face(U[o], bup[o], bdn[o], bup[xr], bdn[xr]);
                V[o] = limit_face(V[o], bup[o], bdn[o], bup[yu], bdn[yu]);
                W[o] = limit_face(W[o], bup[o], bdn[o], bup[zf], bdn[zf]);
            }
}

int orc_per_step(int nx, int ny, int nz, const double* U, const double* V, const double* W,
                 int iord, int nonosc, const double* q, const double* cin, double* cout)
{
    if (nx < 1 || ny < 1 || nz < 1 || iord < 1) return 2;
    size_t n = (size_t)nx * ny * nz;
    int corr = iord > 1;
    double* buf = corr ? malloc(sizeof(double) * n * (nonosc ? 10 : 8)) : NULL;
    if (corr && !buf) return 1;
    double *ca = buf, *cb = buf + n;                        /* pass fields */
    double *ua = buf + 2 * n, *va = buf + 3 * n, *wa = buf + 4 * n;   /* current faces */
    double *ub = buf + 5 * n, *vb = buf + 6 * n, *wb = buf + 7 * n;   /* next faces */
    double *bu = nonosc ? buf + 8 * n : NULL, *bd = nonosc ? buf + 9 * n : NULL;

    if (!corr) {
        orc_per_donor(nx, ny, nz, U, V, W, cin, cout);
    } else {
        const double *cu = U, *cv = V, *cw = W; /* velocities of the previous pass */
        const double* cur = cin;                /* field entering the next pass */
        for (int p = 0; p < iord; p++) {
            double* dst = (p == iord - 1) ? cout : ((p & 1) ? cb : ca);
            if (p == 0) {
                orc_per_donor(nx, ny, nz, U, V, W, cin, dst);
            } else {
                orc_per_antidiff(nx, ny, nz, cur, cu, cv, cw, ub, vb, wb);
                if (nonosc) {
                    orc_per_limiter_beta(nx, ny, nz, cin, cur, ub, vb, wb, bu, bd);
                    orc_per_limiter_apply(nx, ny, nz, bu, bd, ub, vb, wb);
                }
                orc_per_donor(nx, ny, nz, ub, vb, wb, cur, dst);
                double* t;
                t = ua; ua = ub; ub = t;
                t = va; va = vb; vb = t;
                t = wa; wa = wb; wb = t;
                cu = ua; cv = va; cw = wa;
            }
            cur = dst;
        }
    }
    if (q)
        for (size_t o = 0; o < n; o++) cout[o] = cout[o] + q[o];
    free(buf);
    return 0;
}

int orc_per_run(int nx, int ny, int nz, const double* U, const double* V, const double* W,
                int iord, int nonosc, const double* q, double* c, int nsteps)
{
    size_t n = (size_t)nx * ny * nz;
    double* t = malloc(sizeof(double) * n);
    if (!t) return 1;
    for (int s = 0; s < nsteps; s++) {
        int rc = orc_per_step(nx, ny, nz, U, V, W, iord, nonosc, q, c, t);
        if (rc) { free(t); return rc; }
        memcpy(c, t, sizeof(double) * n);
    }
    free(t);
    return 0;
}
