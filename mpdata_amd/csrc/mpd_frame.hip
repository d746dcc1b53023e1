// FRAME mode: the reference step smolar() (core/src/smolar.c:33-63) on the device.
//
// Arrays keep the reference's layouts (cells k*nx*ny + j*nx + i; faces as uu.c:122,
// vv.c:133, ww.c:299) and every formula comes from mpd_formulas.hpp, so the device
// result equals the reference's FP64 result bit for bit.  One thread per face or cell;
// this mode is the drop-in path, the periodic mode carries the tuned kernels.
#include <algorithm>

#include "mpd_formulas.hpp"
#include "mpd_grid.hpp"

namespace mpd {
using namespace exact;
namespace {

constexpr int TPB = 256;

inline unsigned nblocks(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// uu.c:126-138, vv.c:137-149, ww.c:303-315.  blockIdx.y selects the face family.
__global__ void k_frame_courant(int nx, int ny, int ns, const double* __restrict__ dx,
                                const double* __restrict__ dy, const double* __restrict__ ds,
                                double dt, const double* __restrict__ u0,
                                const double* __restrict__ u1, const double* __restrict__ v0,
                                const double* __restrict__ v1, const double* __restrict__ w0,
                                const double* __restrict__ w1, double* __restrict__ U,
                                double* __restrict__ V, double* __restrict__ W)
{
    const size_t np = (size_t)nx * ny;
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (blockIdx.y == 0) {
        size_t n = (size_t)(nx - 1) * (ny - 2) * (ns - 2);
        if (t >= n) return;
        int i = (int)(t % (nx - 1));
        int j = (int)((t / (nx - 1)) % (ny - 2)) + 1;
        int k = (int)(t / ((size_t)(nx - 1) * (ny - 2))) + 1;
        size_t c = k * np + (size_t)j * nx + i;
        double q = 0.25 * dt / dx[i];
        U[t] = q * (((u0[c] + u0[c + 1]) + u1[c]) + u1[c + 1]);
    } else if (blockIdx.y == 1) {
        size_t n = (size_t)(nx - 2) * (ny - 1) * (ns - 2);
        if (t >= n) return;
        int i = (int)(t % (nx - 2)) + 1;
        int j = (int)((t / (nx - 2)) % (ny - 1));
        int k = (int)(t / ((size_t)(nx - 2) * (ny - 1))) + 1;
        size_t c = k * np + (size_t)j * nx + i;
        double q = 0.25 * dt / dy[j];
        V[t] = q * (((v0[c] + v0[c + nx]) + v1[c]) + v1[c + nx]);
    } else {
        size_t n = (size_t)(nx - 2) * (ny - 2) * (ns - 1);
        if (t >= n) return;
        int i = (int)(t % (nx - 2)) + 1;
        int j = (int)((t / (nx - 2)) % (ny - 2)) + 1;
        int k = (int)(t / ((size_t)(nx - 2) * (ny - 2)));
        size_t c = k * np + (size_t)j * nx + i;
        double q = 0.25 * dt / ds[k];
        W[t] = q * (((w0[c] + w0[c + np]) + w1[c]) + w1[c + np]);
    }
}

struct FrameIdx {
    int nx, ny, ns;
    size_t np;
    __device__ size_t C(int i, int j, int k) const { return k * np + (size_t)j * nx + i; }
    __device__ size_t U(int i, int j, int k) const
    {
        return (size_t)(k - 1) * (nx - 1) * (ny - 2) + (size_t)(j - 1) * (nx - 1) + i;
    }
    __device__ size_t V(int i, int j, int k) const
    {
        return (size_t)(k - 1) * (nx - 2) * (ny - 1) + (size_t)j * (nx - 2) + i - 1;
    }
    __device__ size_t W(int i, int j, int k) const
    {
        return (size_t)k * (nx - 2) * (ny - 2) + (size_t)(j - 1) * (nx - 2) + i - 1;
    }
};

// The per-cell and per-face kernels run on 3-D grids of 64 x 4 (x, y) blocks, one z
// plane (or one z plane of one face family) per blockIdx.z: indices come from block
// coordinates, with no 64-bit division (which costs more than the stencil itself).
constexpr int BX = 64, BY = 4;  // a wave per row segment, four rows per workgroup
inline dim3 grid3(int nx, int ny, int nz) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz); }

// smolar.c:51-53 fused: frame cells copy c0 (bndcpy, smolar.c:5-31), interior cells
// take the donor update (smolar_pass.c:30-58).
__global__ void __launch_bounds__(BX * BY)
k_frame_donor(FrameIdx x, int k0, const double* __restrict__ U, const double* __restrict__ V,
              const double* __restrict__ W, const double* __restrict__ c0,
              double* __restrict__ c1)
{
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const int k = k0 + (int)blockIdx.z;
    if (i >= x.nx || j >= x.ny) return;
    const size_t t = x.C(i, j, k);
    if (i == 0 || j == 0 || k == 0 || i == x.nx - 1 || j == x.ny - 1 || k == x.ns - 1) {
        c1[t] = c0[t];
        return;
    }
    c1[t] = donor(c0[t], c0[t - 1], c0[t + 1], c0[t - x.nx], c0[t + x.nx], c0[t - x.np],
                  c0[t + x.np], U[x.U(i - 1, j, k)], U[x.U(i, j, k)], V[x.V(i, j - 1, k)],
                  V[x.V(i, j, k)], W[x.W(i, j, k - 1)], W[x.W(i, j, k)]);
}

// uu.c:56-99, vv.c:58-104, ww.c:222-270: antidiffusive velocities of one corrective
// pass from the previous pass's U0,V0,W0 (all three, pvelocity.c:24-26).  blockIdx.y
// selects the family.  Faces next to the frame take the reduced pattern; the y-faces
// of column i=1 between interior rows are not written (vv.c:81-82).
// Grid: (i, j) over [0, nx-1) x [0, ny-1), blockIdx.z = 3 (k - k0) + family, k in
// [0, ns-1); each family takes the faces of its own index ranges.
__global__ void __launch_bounds__(BX * BY)
k_frame_antidiff(FrameIdx x, int k0, const double* __restrict__ c, const double* __restrict__ U0,
                 const double* __restrict__ V0, const double* __restrict__ W0,
                 double* __restrict__ U1, double* __restrict__ V1, double* __restrict__ W1)
{
    const int nx = x.nx, ny = x.ny, ns = x.ns;
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const int fam = blockIdx.z % 3, k = k0 + (int)(blockIdx.z / 3);
    if (fam == 0) {
        // faces i + 1/2, i in [0, nx-2], of interior rows j, k
        if (i > nx - 2 || j < 1 || j > ny - 2 || k < 1 || k > ns - 2) return;
        const size_t t = x.U(i, j, k);
        double a = U0[t], div, sv, sw;
        if (i == 0 || i == nx - 2) {
            int ii = (i == 0) ? 1 : nx - 2;
            div = 0.0;
            sv = 0.5 * (V0[x.V(ii, j - 1, k)] + V0[x.V(ii, j, k)]);
            sw = 0.5 * (W0[x.W(ii, j, k - 1)] + W0[x.W(ii, j, k)]);
        } else {
            double vr = V0[x.V(i + 1, j - 1, k)], vru = V0[x.V(i + 1, j, k)];
            double vu = V0[x.V(i, j, k)], vc = V0[x.V(i, j - 1, k)];
            double wr = W0[x.W(i + 1, j, k - 1)], wrf = W0[x.W(i + 1, j, k)];
            double wf = W0[x.W(i, j, k)], wc = W0[x.W(i, j, k - 1)];
            sv = 0.25 * (vr + vru + vu + vc);
            sw = 0.25 * (wr + wrf + wf + wc);
            div = 0.25 * a *
                  (U0[t + 1] - U0[t - 1] + vr - vru - vu + vc + wr - wrf - wf + wc);
        }
        size_t o = x.C(i, j, k);
        const size_t np = x.np;
        double cc = c[o], cr = c[o + 1];
        double cru = c[o + nx + 1], cu = c[o + nx], cd = c[o - nx], crd = c[o - nx + 1];
        double crf = c[o + np + 1], cf = c[o + np], cb = c[o - np], crb = c[o - np + 1];
        U1[t] = antidiff(div, a, cr, cc, sv, cru + cu - cd - crd, cru + cu + cd + crd, sw,
                         crf + cf - cb - crb, crf + cf + cb + crb);
    } else if (fam == 1) {
        // faces j + 1/2, j in [0, ny-2], of interior columns i, k
        if (i < 1 || i > nx - 2 || j > ny - 2 || k < 1 || k > ns - 2) return;
        const size_t t = x.V(i, j, k);
        bool edge = (j == 0 || j == ny - 2);
        if (!edge && i == 1) return;  // vv.c:81-82
        double a = V0[t], div, su, sw;
        if (edge) {
            int jj = (j == 0) ? 1 : ny - 2;
            div = 0.0;
            su = 0.5 * (U0[x.U(i - 1, jj, k)] + U0[x.U(i, jj, k)]);
            sw = 0.5 * (W0[x.W(i, jj, k - 1)] + W0[x.W(i, jj, k)]);
        } else {
            double uu = U0[x.U(i - 1, j + 1, k)], uru = U0[x.U(i, j + 1, k)];
            double ur = U0[x.U(i, j, k)], uc = U0[x.U(i - 1, j, k)];
            double ulu = U0[x.U(i - 2, j + 1, k)], ul = U0[x.U(i - 2, j, k)];
            double wu = W0[x.W(i, j + 1, k - 1)], wuf = W0[x.W(i, j + 1, k)];
            double wf = W0[x.W(i, j, k)], wc = W0[x.W(i, j, k - 1)];
            su = 0.25 * (uu + uru + ur + uc);
            sw = 0.25 * (wu + wuf + wf + wc);
            div = 0.25 * a *
                  (V0[t + (nx - 2)] - V0[t - (nx - 2)] + wu - wuf - wf + wc + uu - ulu - ul + uc);
        }
        size_t o = x.C(i, j, k);
        const size_t np = x.np;
        double cc = c[o], cu = c[o + nx];
        double cru = c[o + nx + 1], cr = c[o + 1], cl = c[o - 1], clu = c[o + nx - 1];
        double cuf = c[o + np + nx], cf = c[o + np], cb = c[o - np], cub = c[o - np + nx];
        V1[t] = antidiff(div, a, cu, cc, su, cru + cr - cl - clu, cru + cr + cl + clu, sw,
                         cuf + cf - cb - cub, cuf + cf + cb + cub);
    } else {
        // faces k + 1/2, k in [0, ns-2], of interior columns i, j
        if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2 || k > ns - 2) return;
        const size_t t = x.W(i, j, k);
        double a = W0[t], div, su, sv;
        if (k == 0 || k == ns - 2) {
            int kk = (k == 0) ? 1 : ns - 2;
            div = 0.0;
            su = 0.5 * (U0[x.U(i - 1, j, kk)] + U0[x.U(i, j, kk)]);
            sv = 0.5 * (V0[x.V(i, j - 1, kk)] + V0[x.V(i, j, kk)]);
        } else {
            double ur = U0[x.U(i, j, k)], urf = U0[x.U(i, j, k + 1)];
            double uf = U0[x.U(i - 1, j, k + 1)], uc = U0[x.U(i - 1, j, k)];
            double vf = V0[x.V(i, j - 1, k + 1)], vuf = V0[x.V(i, j, k + 1)];
            double vu = V0[x.V(i, j, k)], vc = V0[x.V(i, j - 1, k)];
            su = 0.25 * (ur + urf + uf + uc);
            sv = 0.25 * (vf + vuf + vu + vc);
            size_t wp = (size_t)(nx - 2) * (ny - 2);
            div = 0.25 * a * (W0[t + wp] - W0[t - wp] + uf - urf - ur + uc + vf - vuf - vu + vc);
        }
        size_t o = x.C(i, j, k);
        const size_t np = x.np;
        double cc = c[o], cf = c[o + np];
        double crf = c[o + np + 1], cr = c[o + 1], cl = c[o - 1], clf = c[o + np - 1];
        double cuf = c[o + np + nx], cu = c[o + nx], cd = c[o - nx], cdf = c[o + np - nx];
        W1[t] = antidiff(div, a, cf, cc, su, crf + cr - cl - clf, crf + cr + cl + clf, sv,
                         cuf + cu - cd - cdf, cuf + cu + cd + cdf);
    }
}

// Periodic face Courant numbers (same formula, wrapped upper neighbour).
__global__ void k_periodic_courant(int nx, int ny, int nz, const double* __restrict__ dx,
                                   const double* __restrict__ dy, const double* __restrict__ dz,
                                   double dt, const double* __restrict__ u0,
                                   const double* __restrict__ u1, const double* __restrict__ v0,
                                   const double* __restrict__ v1, const double* __restrict__ w0,
                                   const double* __restrict__ w1, double* __restrict__ U,
                                   double* __restrict__ V, double* __restrict__ W)
{
    const size_t np = (size_t)nx * ny;
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= np * nz) return;
    int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / np);
    size_t xr = t - i + (i + 1 == nx ? 0 : i + 1);
    size_t yu = t + (size_t)((j + 1 == ny ? 0 : j + 1) - j) * nx;
    size_t zf = t + (size_t)((k + 1 == nz ? 0 : k + 1) - k) * np;
    double qx = 0.25 * dt / dx[i], qy = 0.25 * dt / dy[j], qz = 0.25 * dt / dz[k];
    U[t] = qx * (((u0[t] + u0[xr]) + u1[t]) + u1[xr]);
    V[t] = qy * (((v0[t] + v0[yu]) + v1[t]) + v1[yu]);
    W[t] = qz * (((w0[t] + w0[zf]) + w1[t]) + w1[zf]);
}

}  // namespace

int courant_from_winds(mpd_grid_s* g, const double* dx, const double* dy, const double* dz,
                       double dt, const double* u0, const double* u1, const double* v0,
                       const double* v1, const double* w0, const double* w1)
{
    hipStream_t s = stream();
    if (g->boundary == MPD_FRAME) {
        size_t n = g->nface[0];
        if (g->nface[1] > n) n = g->nface[1];
        if (g->nface[2] > n) n = g->nface[2];
        dim3 grid(nblocks(n), 3);
        hipLaunchKernelGGL(k_frame_courant, grid, dim3(TPB), 0, s, g->nx, g->ny, g->nz, dx,
                           dy, dz, dt, u0, u1, v0, v1, w0, w1, g->face[0].p, g->face[1].p,
                           g->face[2].p);
    } else {
        hipLaunchKernelGGL(k_periodic_courant, dim3(nblocks(g->ncell)), dim3(TPB), 0, s,
                           g->nx, g->ny, g->nz, dx, dy, dz, dt, u0, u1, v0, v1, w0, w1,
                           g->face[0].p, g->face[1].p, g->face[2].p);
    }
    MPD_HIP(hipGetLastError());
    return MPD_OK;
}

// smolar.c:33-63.  The antidiffusive ping-pong of the reference (uu.c:43-44) is kept:
// pass o reads buffer set (o-1)%2 and writes set o%2, where set 0 starts as the
// Courant numbers and set 1 starts zero-filled (the zero-malloc contract; it decides
// the y-face column vv_pass never writes).
int frame_step(mpd_grid_s* g)
{
    hipStream_t s = stream();
    int rc;
    // planes per launch (blockIdx.z <= 65535 with three face families per plane) and
    // rows (blockIdx.y <= 65535)
    constexpr int kMaxPlanes = 65535 / 3;
    if ((g->ny + BY - 1) / BY > 65535)
        return fail(MPD_ERR_UNSUPPORTED, "FRAME grids need ny <= %d", 65535 * BY);
    FrameIdx x{g->nx, g->ny, g->nz, (size_t)g->nx * g->ny};
    double* set[2][3];
    if (g->iord > 1) {
        for (int d = 0; d < 3; d++) {
            if ((rc = g->fb[d].alloc(g->nface[d])) != MPD_OK) return rc;
            MPD_HIP(hipMemsetAsync(g->fb[d].p, 0, g->nface[d] * sizeof(double), s));
            set[1][d] = g->fb[d].p;
            if (g->iord > 2) {  // set 0 is written from pass 2 on: work on a copy
                if ((rc = g->fa[d].alloc(g->nface[d])) != MPD_OK) return rc;
                MPD_HIP(hipMemcpyAsync(g->fa[d].p, g->face[d].p, g->nface[d] * sizeof(double),
                                       hipMemcpyDeviceToDevice, s));
                set[0][d] = g->fa[d].p;
            } else {
                set[0][d] = g->face[d].p;
            }
        }
    } else {
        for (int d = 0; d < 3; d++) set[0][d] = g->face[d].p;
    }
    int cold = g->cur;
    for (int o = 0; o < g->iord; o++) {
        double** vel = set[o & 1];
        if (o) {
            double** prev = set[(o - 1) & 1];
            for (int k0 = 0; k0 < g->nz - 1; k0 += kMaxPlanes)
                hipLaunchKernelGGL(k_frame_antidiff,
                                   grid3(g->nx - 1, g->ny - 1, 3 * std::min(kMaxPlanes, g->nz - 1 - k0)),
                                   dim3(BX, BY), 0, s, x, k0, g->c[cold].p, prev[0], prev[1],
                                   prev[2], vel[0], vel[1], vel[2]);
        }
        for (int k0 = 0; k0 < g->nz; k0 += kMaxPlanes)
            hipLaunchKernelGGL(k_frame_donor, grid3(g->nx, g->ny, std::min(kMaxPlanes, g->nz - k0)),
                               dim3(BX, BY), 0, s, x, k0, vel[0], vel[1], vel[2], g->c[cold].p,
                               g->c[1 - cold].p);
        cold = 1 - cold;
    }
    MPD_HIP(hipGetLastError());
    g->cur = cold;
    return MPD_OK;
}

}  // namespace mpd
