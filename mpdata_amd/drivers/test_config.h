/* The test-configuration interface the reference's driver programs are written against
 * (tests/init.h:1-44: struct test_config_t and the test_* functions), implemented on
 * this project's driver set-up (ring_setup.c), so that the reference's unchanged
 * tests/ring2d/test.c and tests/ring3d/test.c build against include/smolar.h and link
 * against libsmolar_hip.so.  The struct's members and their order are the reference's
 * (callers access them directly).  VAPOR is not available here: the two VAPOR
 * functions write raw dumps instead. */
#pragma once
#include "../../include/smolar.h"

struct test_config_t {
    void* desc;
    int n, m;
    real dt;
    real *u, *v, *w;
    real *c0, *c1;
    real *dx, *dy, *ds;
    int order, flow_type;
    int xof, yof;
};

/* The configuration of tests/init.c:15-103 on an n x n x m grid, in ONE allocation (the
 * callers release it with free(t)); *size = its byte size.  NULL when out of memory. */
struct test_config_t* test_init(int n, int m, int* size);
/* A copy of t (size bytes, from test_init) with its pointers into the copy. */
struct test_config_t* test_clone(int size, struct test_config_t* t);
/* GrADS descriptor, script and launcher of <name>.bin (tests/init.c:124-203). */
void test_create_grads_ctl(int n, int nt, const char* name);
void test_create_grads_gs(int n, int nt, const char* name);
void test_create_grads_pl(const char* name);
/* In place of the reference's VAPOR dataset: <name>.raw, truncated here, to which every
 * test_write_vapor_vdf call appends u, v, w and c (n*n*m doubles each). */
void test_create_vapor_vdf(int n, int m, int nt, const char* name);
void test_write_vapor_vdf(int n, int m, int it, const char* name, real* u, real* v, real* w,
                          real* c);
