/* ring3d <n> <nt>: the reference's 3-D ring test (tests/ring3d/test.c:10-52) on the
 * HIP-backed smolar().  The reference writes VAPOR datasets; this driver writes
 * ring3d.bin instead: u, v, w and the initial c (n^3 doubles each), then c after every
 * step. */
#include <stdio.h>
#include <stdlib.h>

#include "../../include/smolar.h"
#include "../../include/timing.h"
#include "ring_setup.h"

int main(int argc, char* argv[])
{
    const char* name = "ring3d";
    if (argc != 3) {
        printf("%s <n> <nt>\n", name);
        return 0;
    }
    int n = atoi(argv[1]), nt = atoi(argv[2]);
    if (n < 4 || nt < 0) {
        fprintf(stderr, "n must be >= 4 and nt >= 0\n");
        return 1;
    }
    size_t nn = (size_t)n * n * n;
    struct ring_config* t = ring_init(n, n);
    if (!t) {
        fprintf(stderr, "out of memory\n");
        return 1;
    }
    char fn[64];
    snprintf(fn, sizeof fn, "%s.bin", name);
    FILE* fp = fopen(fn, "wb");
    if (!fp) {
        perror(fn);
        return 1;
    }
    fwrite(t->u, 8, nn, fp);
    fwrite(t->v, 8, nn, fp);
    fwrite(t->w, 8, nn, fp);
    fwrite(t->c0, 8, nn, fp);
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
        fwrite(t->c1, 8, nn, fp);
        printf("step %d\n", it);
        double* c = t->c0;
        t->c0 = t->c1;
        t->c1 = c;
    }
    smolar_get_time(&b);
    fclose(fp);
    fprintf(stderr, "%d steps in %.6f s\n", nt, smolar_get_time_diff(&a, &b));
    ring_free(t);
    return 0;
}
