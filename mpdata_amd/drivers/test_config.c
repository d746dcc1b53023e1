/* The reference's test-configuration interface (tests/init.h) on this project's driver
 * set-up; see test_config.h. */
#include "test_config.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ring_setup.h"

/* Byte offsets of the arrays after the struct in the single block. */
struct layout {
    size_t u, v, w, c0, c1, dx, dy, ds, total;
};

static struct layout layout_of(int n, int m)
{
    struct layout L;
    size_t nn = (size_t)n * n * m * sizeof(real);
    size_t off = (sizeof(struct test_config_t) + 63) / 64 * 64;
    L.u = off;
    L.v = L.u + nn;
    L.w = L.v + nn;
    L.c0 = L.w + nn;
    L.c1 = L.c0 + nn;
    L.dx = L.c1 + nn;
    L.dy = L.dx + n * sizeof(real);
    L.ds = L.dy + n * sizeof(real);
    L.total = L.ds + m * sizeof(real);
    return L;
}

static void point(struct test_config_t* t, const struct layout* L)
{
    char* b = (char*)t;
    t->u = (real*)(b + L->u);
    t->v = (real*)(b + L->v);
    t->w = (real*)(b + L->w);
    t->c0 = (real*)(b + L->c0);
    t->c1 = (real*)(b + L->c1);
    t->dx = (real*)(b + L->dx);
    t->dy = (real*)(b + L->dy);
    t->ds = (real*)(b + L->ds);
}

struct test_config_t* test_init(int n, int m, int* size)
{
    struct ring_config* r = ring_init(n, m);
    if (!r) return NULL;
    struct layout L = layout_of(n, m);
    if (L.total > 0x7fffffff) {
        ring_free(r);
        return NULL;
    }
    struct test_config_t* t = calloc(1, L.total);
    if (!t) {
        ring_free(r);
        return NULL;
    }
    point(t, &L);
    size_t nn = (size_t)n * n * m * sizeof(real);
    t->desc = NULL;
    t->n = n;
    t->m = m;
    t->dt = r->dt;
    memcpy(t->u, r->u, nn);
    memcpy(t->v, r->v, nn);
    memcpy(t->w, r->w, nn);
    memcpy(t->c0, r->c0, nn);
    memcpy(t->c1, r->c1, nn);
    memcpy(t->dx, r->dx, n * sizeof(real));
    memcpy(t->dy, r->dy, n * sizeof(real));
    memcpy(t->ds, r->ds, m * sizeof(real));
    t->order = r->order;
    t->flow_type = r->flow_type;
    t->xof = r->xof;
    t->yof = r->yof;
    ring_free(r);
    if (size) *size = (int)L.total;
    return t;
}

struct test_config_t* test_clone(int size, struct test_config_t* t)
{
    struct test_config_t* c = malloc((size_t)size);
    if (!c) return NULL;
    memcpy(c, t, (size_t)size);
    struct layout L = layout_of(t->n, t->m);
    point(c, &L);
    /* the caller may have swapped c0 and c1: keep which array each points to */
    if (t->c0 == (real*)((char*)t + L.c1)) {
        real* x = c->c0;
        c->c0 = c->c1;
        c->c1 = x;
    }
    return c;
}

void test_create_grads_ctl(int n, int nt, const char* name)
{
    (void)nt; /* the descriptor's time axis is fixed (tests/init.c:131) */
    ring_write_grads_ctl(n, name);
}

void test_create_grads_gs(int n, int nt, const char* name) { ring_write_grads_gs(n, nt, name); }

void test_create_grads_pl(const char* name) { ring_write_grads_pl(name); }

void test_create_vapor_vdf(int n, int m, int nt, const char* name)
{
    (void)n;
    (void)m;
    (void)nt;
    char fn[256];
    snprintf(fn, sizeof fn, "%s.raw", name);
    FILE* fp = fopen(fn, "wb");
    if (fp) fclose(fp);
}

void test_write_vapor_vdf(int n, int m, int it, const char* name, real* u, real* v, real* w,
                          real* c)
{
    (void)it;
    char fn[256];
    snprintf(fn, sizeof fn, "%s.raw", name);
    FILE* fp = fopen(fn, "ab");
    if (!fp) return;
    size_t nn = (size_t)n * n * m;
    fwrite(u, sizeof(real), nn, fp);
    fwrite(v, sizeof(real), nn, fp);
    fwrite(w, sizeof(real), nn, fp);
    fwrite(c, sizeof(real), nn, fp);
    fclose(fp);
}
