/* ring2d <n> <nt>: the reference's 2-D ring test (tests/ring2d/test.c:10-71) on the
 * HIP-backed smolar().  Writes ring2d.bin (per step: the u, v and c planes of layer 1,
 * n*n doubles each; first the initial state) with its GrADS descriptor, script and perl
 * launcher (ring2d.ctl, ring2d.gs, ring2d.pl, the ring2d_frames directory), and prints
 * the probe cell each step. */
#include <stdio.h>
#include <stdlib.h>

#include "../../include/smolar.h"
#include "../../include/timing.h"
#include "ring_setup.h"

int main(int argc, char* argv[])
{
    const char* name = "ring2d";
    if (argc != 3) {
        printf("%s <n> <nt>\n", name);
        return 0;
    }
    int n = atoi(argv[1]), nt = atoi(argv[2]);
    if (n < 3 || nt < 0) {
        fprintf(stderr, "n must be >= 3 and nt >= 0\n");
        return 1;
    }
    size_t n2 = (size_t)n * n;
    struct ring_config* t = ring_init(n, 3);
    if (!t) {
        fprintf(stderr, "out of memory\n");
        return 1;
    }
    ring_write_grads_ctl(n, name); /* ring2d/test.c:26-33 */
    ring_write_grads_gs(n, nt, name);
    ring_write_grads_pl(name);
    char fn[64];
    snprintf(fn, sizeof fn, "%s.bin", name);
    FILE* fp = fopen(fn, "wb");
    if (!fp) {
        perror(fn);
        return 1;
    }
    fwrite(t->u + n2, 8, n2, fp);
    fwrite(t->v + n2, 8, n2, fp);
    fwrite(t->c0 + n2, 8, n2, fp);
    printf("step 0\n");
    struct smolar_time_t a, b;
    smolar_get_time(&a);
    for (int it = 1; it <= nt; it++) {
        int rc = smolar(t->n, t->n, t->m, t->dx, t->dy, t->ds, t->dt, t->u, t->u, t->v, t->v,
                        t->w, t->w, &t->c0, &t->c1, t->flow_type, t->order);
        if (rc) {
            fprintf(stderr, "smolar failed with code %d\n", rc);
            return 1;
        }
        fwrite(t->u + n2, 8, n2, fp);
        fwrite(t->v + n2, 8, n2, fp);
        fwrite(t->c1 + n2, 8, n2, fp);
        printf("step %d %f\n", it, t->c1[n / 2 + t->xof + n * (n / 2 + t->yof) + n2]);
        double* c = t->c0;  /* ring2d/test.c:62-63 */
        t->c0 = t->c1;
        t->c1 = c;
    }
    smolar_get_time(&b);
    fclose(fp);
    fprintf(stderr, "%d steps in %.6f s\n", nt, smolar_get_time_diff(&a, &b));
    ring_free(t);
    return 0;
}
