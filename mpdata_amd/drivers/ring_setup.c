/* Set-up of the ring2d / ring3d drivers (tests/init.c:15-103 and the generators it calls,
 * generators/src/genring2d.c:8-64, genringf2d.c:9-67, gengauss2d.c:6-33,
 * gengauss3d.c:6-39), restated so the drivers do not need the reference's sources. */
#include "ring_setup.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

/* Ring wind (genring2d.c:29-44 and genringf2d.c:31-46, the non-uniform variants):
 * solid-body rotation scaled by 1/rmin inside rmin <= r <= rmax, faded by exp(rmax - r)
 * outside; zero inside rmin. */
static void ring_wind(int cx, int cy, int nx, int ny, const double* dx, const double* dy,
                      double rmin, double rmax, double* u, double* v)
{
    for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++) {
            double x = (double)(ix - cx), y = (double)(iy - cy);
            double r = sqrt(x * x + y * y);
            if (r >= rmin && r <= rmax) {
                u[ix + iy * nx] = -y / rmin * dx[ix];
                v[ix + iy * nx] = x / rmin * dy[iy];
            }
        }
    for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++) {
            double x = (double)(ix - cx), y = (double)(iy - cy);
            double r = sqrt(x * x + y * y);
            if (r > rmax) {
                u[ix + nx * iy] = -y / rmin * dx[ix] * exp((rmax - r) * dx[ix]);
                v[ix + nx * iy] = x / rmin * dy[iy] * exp((rmax - r) * dy[iy]);
            }
        }
}

/* Gaussian clipped at `radius` (gengauss2d.c:14-23, gengauss3d.c:18-31). */
static void gauss(int cx, int cy, int cs, int nx, int ny, int ns, double radius, double sigma,
                  double* g)
{
    double s2 = sigma * sigma;
    for (int is = 0; is < ns; is++)
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int a = ix - cx, b = iy - cy, c = is - cs;
                double a2 = a * a, b2 = b * b, c2 = c * c;
                if (sqrt(a2 + b2 + c2) <= radius)
                    g[ix + nx * iy + (size_t)nx * ny * is] =
                        exp(-0.5 * (a2 / s2 + b2 / s2 + c2 / s2));
            }
}

static void gauss2d(int cx, int cy, int nx, int ny, double radius, double sigma, double* g)
{
    double s2 = sigma * sigma;
    for (int iy = 0; iy < ny; iy++)
        for (int ix = 0; ix < nx; ix++) {
            int a = ix - cx, b = iy - cy;
            double a2 = a * a, b2 = b * b;
            if (sqrt(a2 + b2) <= radius)
                g[ix + nx * iy] = exp(-0.5 * (a2 / s2 + b2 / s2));
        }
}

struct ring_config* ring_init(int n, int m)
{
    struct ring_config* t = calloc(1, sizeof *t);
    if (!t) return NULL;
    size_t n2 = (size_t)n * n, nn = n2 * m;
    t->n = n;
    t->m = m;
    t->dt = 0.0125;           /* init.c:22 */
    t->order = 3;             /* init.c:29 */
    t->flow_type = 1;
    t->xof = n / 4;           /* init.c:25 */
    t->yof = 0;
    double rmin = 1, rmax = (double)(float)(n / 2.5f); /* init.c:19: a float expression */
    double sigma = 6.0, radius = (double)(n / 8);      /* init.c:26 */
    t->u = calloc(nn, 8); t->v = calloc(nn, 8); t->w = calloc(nn, 8);
    t->c0 = calloc(nn, 8); t->c1 = calloc(nn, 8);
    t->dx = malloc(8 * n); t->dy = malloc(8 * n); t->ds = malloc(8 * m);
    if (!t->u || !t->v || !t->w || !t->c0 || !t->c1 || !t->dx || !t->dy || !t->ds) {
        ring_free(t);
        return NULL;
    }
    for (int i = 0; i < m; i++) t->ds[i] = 1;
    for (int i = 0; i < n; i++) t->dx[i] = t->dy[i] = 1.0;
    if (m > 3)  /* init.c:64-68: 3-D spiral, constant vertical speed */
        for (size_t i = 0; i < nn; i++) t->w[i] = 1.0;
    ring_wind(n / 2, n / 2, n, n, t->dx, t->dy, rmin, rmax, t->u, t->v);
    for (int i = 1; i < m; i++) {
        memcpy(t->u + i * n2, t->u, n2 * 8);
        memcpy(t->v + i * n2, t->v, n2 * 8);
    }
    if (m > 3) {
        gauss(n / 2 + t->xof, n / 2 + t->yof, (int)radius, n, n, m, radius, sigma, t->c0);
    } else {
        gauss2d(n / 2 + t->xof, n / 2 + t->yof, n, n, radius, sigma, t->c0);
        for (int i = 1; i < m; i++) memcpy(t->c0 + i * n2, t->c0, n2 * 8);
    }
    return t;
}

void ring_free(struct ring_config* t)
{
    if (!t) return;
    free(t->u); free(t->v); free(t->w); free(t->c0); free(t->c1);
    free(t->dx); free(t->dy); free(t->ds);
    free(t);
}

void ring_write_grads_ctl(int n, const char* name)
{
    char fn[256];
    snprintf(fn, sizeof fn, "%s.ctl", name);
    FILE* fp = fopen(fn, "w");
    if (!fp) return;
    fprintf(fp, "DSET ^%s.bin\n", name);
    fprintf(fp, "XDEF %d LINEAR 0 1\n", n);
    fprintf(fp, "YDEF %d LINEAR 0 1\n", n);
    fprintf(fp, "ZDEF 1 LINEAR 0 1\n");
    fprintf(fp, "EDEF 3 NAMES u v c\n");
    fprintf(fp, "TDEF 4001 LINEAR 18:00Z04jul2000 1hr\n");
    fprintf(fp, "UNDEF 1E+300\n");
    fprintf(fp, "VARS 3\n");
    fprintf(fp, "u 1 -1 uwind\n");
    fprintf(fp, "v 1 -1 vwind\n");
    fprintf(fp, "c 1 -1 source\n");
    fprintf(fp, "ENDVARS\n");
    fclose(fp);
}

/* GrADS script drawing each of the nt stored steps into <name>_frames/stepNNNN.png:
 * shaded c with contour levels 0, 0.1, .., 1 and the (u, v) vectors.  The text is the
 * reference's, byte for byte, including the missing blank between the 0.4 and 0.5
 * levels of its `set clevs` line. */
void ring_write_grads_gs(int n, int nt, const char* name)
{
    char fn[256];
    snprintf(fn, sizeof fn, "%s.gs", name);
    FILE* fp = fopen(fn, "w");
    if (!fp) return;
    fprintf(fp, "'open '\"%s.ctl\"\n", name);
    for (int it = 0; it < nt; it++) {
        fputs("'set parea '0.5' '10.5' '1.5' '7.5\n", fp);
        for (int axis = 0; axis < 2; axis++)
            fprintf(fp, "'set %caxis '0' '%d' '16\n", "xy"[axis], n);
        fputs("'set mproj 'off\n'set mpdraw 'off\n'set gxout 'shaded\n", fp);
        fputs("'set clevs ", fp);
        for (int l = 0; l <= 10; l++) /* levels 0.0 .. 1.0; none after the last */
            fprintf(fp, "'%f%s", 0.1 * l, l == 4 ? "'" : (l < 10 ? "' " : "\n"));
        fprintf(fp, "'set t '%d\n'display 'c\n", it + 1);
        fprintf(fp, "'draw title '\"Source in %s flow - step %04d\"\n", name, it);
        fputs("'run 'cbarm.gs\n'set gxout 'vector\n", fp);
        fputs("'display '\"skip(u,4); skip(v,4); mag(u,v)\"\n", fp);
        fprintf(fp, "'printim '\"%s_frames/step%04d.png\"' 'png' 'x800' 'y1024\n", name, it);
        fputs("'clear 'graphics\n", fp);
    }
    fclose(fp);
}

/* Perl launcher of the GrADS script, and the <name>_frames directory it draws into. */
void ring_write_grads_pl(const char* name)
{
    char fn[256];
    snprintf(fn, sizeof fn, "%s_frames", name);
    mkdir(fn, 0777);
    snprintf(fn, sizeof fn, "%s.pl", name);
    FILE* fp = fopen(fn, "w");
    if (!fp) return;
    fprintf(fp, "#!/usr/bin/perl -w\nsystem(\"grads -lbxc \\\"run %s.gs\\\"\");\n", name);
    fclose(fp);
}
