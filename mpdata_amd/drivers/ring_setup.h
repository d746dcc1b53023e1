/* Set-up of the ring2d / ring3d drivers: the reference test configuration
 * (tests/init.c:15-103) and its field generators, restated. */
#pragma once
#include "../../include/smolar.h"

struct ring_config {
    int n, m;           /* grid n x n x m (m = 3: the 2-D embedding) */
    double dt;
    int order, flow_type;
    int xof, yof;       /* blob offset from the centre */
    double *u, *v, *w;  /* cell-centred winds, n*n*m each */
    double *c0, *c1;    /* field and scratch, n*n*m each */
    double *dx, *dy, *ds;
};

/* tests/init.c:15-103; returns NULL when out of memory.  Free with ring_free. */
struct ring_config* ring_init(int n, int m);
void ring_free(struct ring_config* t);
/* tests/init.c:124-145 (GrADS descriptor of <name>.bin) */
void ring_write_grads_ctl(int n, const char* name);
/* tests/init.c:148-183 (GrADS script drawing the nt steps of <name>.bin) */
void ring_write_grads_gs(int n, int nt, const char* name);
/* tests/init.c:186-203 (perl launcher of the script; creates <name>_frames) */
void ring_write_grads_pl(const char* name);
