/* Drop-in C interface of the MPDATA transport step, backed by HIP kernels on an AMD
 * Instinct MI355X (gfx950).  Source-compatible with the reference header smolar.h:
 * the same symbols, argument meaning and pointer contract, with real = double
 * (the library computes in FP64 only).
 *
 * Link: -lsmolar_hip (mpdata_amd/libsmolar_hip.so).
 */
#pragma once

#define SOLENOIDAL 0 /* smolar.h:3 */
#define DIVERGENT 1  /* smolar.h:4 */

#if defined(SINGLE) || defined(INTEGER)
#error "libsmolar_hip computes in FP64 only: build callers with -DDOUBLE (or no precision flag)"
#endif
#ifndef real
#define real double
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* One MPDATA step on an nx*ny*ns grid whose outer one-cell frame is held fixed.
 * Replaces smolar() (reference smolar.h:45-48, core/src/smolar.c:33-63).
 *   dx[nx], dy[ny], ds[ns]  grid spacings; u0..w1 cell-centred winds at the previous
 *   and current time levels (nx*ny*ns each, i fastest); *c0 the field at time n.
 *   order = number of passes (1 = donor cell only, 2.. = MPDATA corrective passes).
 * On return *c1 points to the new field and *c0 to the previous pass's field; for even
 * order the two pointers are swapped, exactly as the reference does (smolar.c:58).
 * flow_type is accepted and ignored, as in the reference (uu.c:111).
 * All pointers are host pointers; the step runs on the current HIP device.
 * Returns 0 on success, or an MPD_ERR_* code from mpdata.h (the reference always
 * returns 0); mpd_error_string() describes the failure. */
int smolar(int nx, int ny, int ns,
           real* dx, real* dy, real* ds, const real dt,
           real* u0, real* u1, real* v0, real* v1, real* w0, real* w1,
           real** c0, real** c1, int flow_type, int order);

#define MSMOLAR_SERIAL 0   /* smolar.h:53 */
#define MSMOLAR_PTHREADS 1 /* smolar.h:58 */
#define MSMOLAR_OPENMP 2   /* smolar.h:63 */

/* Stateful step.  Replaces msmolar() (smolar.h:95-99, core/src/msmolar.c:13-188).
 * The first call (*desc == NULL) uploads the spacings, winds and field into a
 * device-resident context and stores its handle in *desc; like the reference, later
 * calls ignore the array arguments and advance the state kept in *desc.  sx, sy, ss
 * and mode (the reference's host decomposition and threading hints) are accepted and
 * ignored: the grid is tiled on the device.  Results stay on the device until
 * msmolar_gather(). */
int msmolar(void** desc, int nx, int ny, int ns,
            real* dx, real* dy, real* ds, const real dt,
            real* u0, real* u1, real* v0, real* v1, real* w0, real* w1,
            real** c0, real** c1, int flow_type, int order,
            int sx, int sy, int ss, int mode);

/* Copies the state of *desc back to the caller: the newest field to *c1 and the
 * previous pass's field to *c0 (msmolar_gather.c:17-27).  Replaces msmolar_gather()
 * (smolar.h:132-136). */
int msmolar_gather(void** desc, int nx, int ny, int ns,
                   real* dx, real* dy, real* ds, const real dt,
                   real* u0, real* u1, real* v0, real* v1, real* w0, real* w1,
                   real** c0, real** c1, int flow_type, int order,
                   int sx, int sy, int ss, int mode);

/* Releases a context made by msmolar() and sets *desc to NULL.  No counterpart in the
 * reference, which never frees its descriptor. */
int msmolar_free(void** desc);

#ifdef __cplusplus
}
#endif
