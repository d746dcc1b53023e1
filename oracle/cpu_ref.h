/* CPU oracle for the MPDATA transport step -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load this
 * library, and only as the checker.  The product (mpdata_amd/libsmolar_hip.so) never
 * links, loads or falls back to it.
 *
 * Two families of functions:
 *
 *  FRAME  -- a restatement of the reference step smolar() (core/src/smolar.c:33-63) in
 *            the reference's own array layouts, reproducing its arithmetic order so that
 *            results agree bit for bit with the reference compiled -DDOUBLE
 *            -ffp-contract=off.  Pinned by the golden fixtures under tests/golden/
 *            (made from the compiled reference) and, where the reference tree is
 *            mounted, by direct comparison with oracle/_ref/libsmolar_ref.so.
 *
 *  PER    -- the doubly/triply periodic step used by every benchmark configuration.
 *            The antidiffusive-velocity and donor formulas are the reference's interior
 *            formulas applied with wrap-around indexing.  Pinned against the reference
 *            by the padded-halo construction (reference smolar() on a grid padded by H
 *            periodic cells, central region compared bit for bit).  The nonoscillatory
 *            limiter and the emission term have no counterpart in the reference; they
 *            are pinned by invariants and closed forms (tests/test_oracle_periodic.py).
 *
 * Layout: every cell array is c[k*nx*ny + j*nx + i] (i fastest).  PER face arrays use the
 * same layout: U[idx(i,j,k)] is the x-face between cells i and i+1 (mod nx), V the
 * y-face between j and j+1, W the z-face between k and k+1.
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- FRAME mode: reference layouts ---------------------------------
 * U: (nx-1)(ny-2)(ns-2), index (k-1)(nx-1)(ny-2) + (j-1)(nx-1) + i,  i=0..nx-2
 * V: (nx-2)(ny-1)(ns-2), index (k-1)(nx-2)(ny-1) + j(nx-2) + i-1,    j=0..ny-2
 * W: (nx-2)(ny-2)(ns-1), index k(nx-2)(ny-2) + (j-1)(nx-2) + i-1,    k=0..ns-2   */
void orc_frame_courant(int nx, int ny, int ns,
                       const double* dx, const double* dy, const double* ds, double dt,
                       const double* u0, const double* u1, const double* v0,
                       const double* v1, const double* w0, const double* w1,
                       double* U, double* V, double* W);
void orc_frame_bndcpy(int nx, int ny, int ns, double* to, const double* from);
void orc_frame_donor(int nx, int ny, int ns, const double* U, const double* V,
                     const double* W, const double* c0, double* c1);
void orc_frame_antidiff_u(int nx, int ny, int ns, const double* c, const double* U0,
                          const double* V0, const double* W0, double* U1);
void orc_frame_antidiff_v(int nx, int ny, int ns, const double* c, const double* U0,
                          const double* V0, const double* W0, double* V1);
void orc_frame_antidiff_w(int nx, int ny, int ns, const double* c, const double* U0,
                          const double* V0, const double* W0, double* W1);
/* Same signature and pointer contract as the reference smolar() (smolar.h:45-48). */
int orc_smolar(int nx, int ny, int ns, double* dx, double* dy, double* ds, const double dt,
               double* u0, double* u1, double* v0, double* v1, double* w0, double* w1,
               double** c0, double** c1, int flow_type, int order);

/* ---------------- PER mode: periodic, all arrays nx*ny*nz ------------------------ */
void orc_per_courant(int nx, int ny, int nz,
                     const double* dx, const double* dy, const double* dz, double dt,
                     const double* u0, const double* u1, const double* v0,
                     const double* v1, const double* w0, const double* w1,
                     double* U, double* V, double* W);
void orc_per_donor(int nx, int ny, int nz, const double* U, const double* V,
                   const double* W, const double* cin, double* cout);
void orc_per_antidiff(int nx, int ny, int nz, const double* c, const double* U0,
                      const double* V0, const double* W0, double* U1, double* V1,
                      double* W1);
/* Limiter bounds: bup/bdn per cell from c^n (cn), the pass input (cp) and the
 * unlimited antidiffusive velocities. */
void orc_per_limiter_beta(int nx, int ny, int nz, const double* cn, const double* cp,
                          const double* U, const double* V, const double* W,
                          double* bup, double* bdn);
/* Replace U,V,W by their limited values using bup/bdn. */
void orc_per_limiter_apply(int nx, int ny, int nz, const double* bup, const double* bdn,
                           double* U, double* V, double* W);
/* One full step: donor pass, then iord-1 corrective passes (each optionally limited),
 * then cout += q when q is non-NULL.  U,V,W are the face Courant numbers. */
int orc_per_step(int nx, int ny, int nz, const double* U, const double* V, const double* W,
                 int iord, int nonosc, const double* q, const double* cin, double* cout);
/* Threads of the periodic loops (OpenMP over z planes; results do not depend on it). */
void orc_set_threads(int n);
int orc_get_threads(void);
/* nsteps steps in place on c. */
int orc_per_run(int nx, int ny, int nz, const double* U, const double* V, const double* W,
                int iord, int nonosc, const double* q, double* c, int nsteps);

/* ---------------- derivatives of one periodic step (iord = 2), cpu_adj.c ---------
 * c_out = A(c_in) + q with fixed face Courant numbers; nonosc selects the limiter.
 * TLM:     dout = A'(c_in) dc_in + dq           (dq may be NULL)
 * adjoint: lin = A'(c_in)^T lout; lq += lout     (lq may be NULL) */
int orc_tlm_step(int nx, int ny, int nz, const double* U, const double* V, const double* W,
                 int nonosc, const double* c0, const double* dc0, const double* dq,
                 double* dout);
int orc_adj_step(int nx, int ny, int nz, const double* U, const double* V, const double* W,
                 int nonosc, const double* c0, const double* lout, double* lin,
                 double* lq);

#ifdef __cplusplus
}
#endif
