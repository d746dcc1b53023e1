// FRAME mode: the reference step smolar() (core/src/smolar.c:33-63) on the device.
//
// Arrays keep the reference's layouts (cells k*nx*ny + j*nx + i; faces as uu.c:122,
// vv.c:133, ww.c:299) and every formula comes from mpd_formulas.hpp, so the device
// result equals the reference's FP64 result bit for bit.  One thread per face or cell;
// this mode is the drop-in path, the periodic mode carries the tuned kernels.
#include <algorithm>

#include "mpd_formulas.hpp"
#include "mpd_grid.hpp"
#include "mpd_tile.hpp"

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

// ------------------------------------------- donor + antidiffusive velocities, tiled ---
//
// One corrective pass's antidiffusive velocities -- and on the first pass the donor
// field they are formed from -- in one LDS-tiled 2.5-D streaming kernel: the frame-mode
// counterpart of the periodic fused F1 (mpd_fused.hip), on the reference's layouts and
// with its frame rules.  Same values as k_frame_donor followed by k_frame_antidiff, bit
// for bit (the same formulas in the same order from mpd_formulas.hpp).
//
// A workgroup owns a TX x TY column tile of the grid (frame included) and a chunk of
// z-planes; each thread owns a tile column and at most one column of the 2-cell halo
// (tile::Geo).  Iteration p: publish c^n(p), U0(p), V0(p) -> barrier -> donor field
// c1(p) on the tile and ring 1 (frame cells: c^n), publish it with W0(p-1) -> barrier ->
// the velocities of the faces above / right of the tile cells of plane k = p - 1.
// Cells and faces outside the grid, or absent from the reference's layouts, read as 0
// and are never used by a face or cell that exists.
struct FrameF1Args {
    int nx, ny, ns, tiles_x, tiles_y, zc;
    const double* cin;            // first pass: c^n; later passes: the pass input
    const double *U0, *V0, *W0;   // velocities of the previous pass (reference layouts)
    double* c1out;                // first pass: the donor field, every cell
    double *U1, *V1, *W1;         // this pass's velocities
    // the y-faces of column i = 1 between interior rows, which vv_pass never writes
    // (vv.c:81-82), receive what the reference's buffer holds there: zero on odd passes
    // (the zero-filled second buffer), the first-order V on even ones (the buffer that
    // starts as the Courant numbers) -- q7 = null or the Courant V array
    const double* q7;
};

// Byte offsets of a cell (gx, gy) within a plane of each layout (32-bit: the host keeps
// every plane below 2^29 cells) and whether the column exists there.  Accesses are a
// uniform plane base plus these offsets (the scalar-base addressing form).
struct FrameCol {
    bool c, u, v, w;
    unsigned oc, ou, ov, ow;
    __device__ FrameCol(const FrameF1Args& a, int gx, int gy)
    {
        const int nx = a.nx, ny = a.ny;
        c = gx >= 0 && gx < nx && gy >= 0 && gy < ny;
        u = gx >= 0 && gx <= nx - 2 && gy >= 1 && gy <= ny - 2;
        v = gx >= 1 && gx <= nx - 2 && gy >= 0 && gy <= ny - 2;
        w = gx >= 1 && gx <= nx - 2 && gy >= 1 && gy <= ny - 2;
        oc = c ? (unsigned)(gy * nx + gx) * 8u : 0u;
        ou = u ? (unsigned)((gy - 1) * (nx - 1) + gx) * 8u : 0u;
        ov = v ? (unsigned)(gy * (nx - 2) + gx - 1) * 8u : 0u;
        ow = w ? (unsigned)((gy - 1) * (nx - 2) + gx - 1) * 8u : 0u;
    }
    __device__ void keep()
    {
        tile::keep32(oc);
        tile::keep32(ou);
        tile::keep32(ov);
        tile::keep32(ow);
    }
};

template <int TX, int TY, bool DONOR>
__global__ void __launch_bounds__((tile::Geo<TX, TY>::NT)) k_frame_f1(FrameF1Args a)
{
    using G = tile::Geo<TX, TY>;
    constexpr int RX = G::RX, R = G::R;
    __shared__ double s_c0[R], s_c1[3][R], s_u[3][R], s_v[3][R], s_w[R];
    const int nx = a.nx, ny = a.ny, ns = a.ns;
    const size_t np = (size_t)nx * ny;
    const size_t upl = (size_t)(nx - 1) * (ny - 2), vpl = (size_t)(nx - 2) * (ny - 1);
    const size_t wpl = (size_t)(nx - 2) * (ny - 2);  // elements per plane of U, V, W

    const int t = threadIdx.x;
    const int item = tile::xcd_item(blockIdx.x, gridDim.x);
    const int bx = item % a.tiles_x, by = (item / a.tiles_x) % a.tiles_y;
    const int bz = item / (a.tiles_x * a.tiles_y);
    const int rx = 2 + t % TX, ry = 2 + t / TX;
    const int L = ry * RX + rx;
    const int gx = bx * TX + rx - 2, gy = by * TY + ry - 2;
    FrameCol P(a, gx, gy);
    const bool hl = t < G::NH, h1 = t < G::H1;
    int hx = 0, hy = 0;
    if (hl) G::halo(t, hx, hy);
    const int HL = hy * RX + hx;
    const int hgx = bx * TX + hx - 2, hgy = by * TY + hy - 2;
    FrameCol H(a, hl ? hgx : -1, hl ? hgy : -1);
    const bool frame_xy = gx == 0 || gx == nx - 1 || gy == 0 || gy == ny - 1;
    const bool hframe_xy = hgx == 0 || hgx == nx - 1 || hgy == 0 || hgy == ny - 1;

    // plane loads: zero outside the grid or the layout
    auto ldc = [&](const FrameCol& q, int k) {
        return (q.c && k >= 0 && k < ns) ? tile::ld(a.cin + (size_t)k * np, q.oc) : 0.0;
    };
    auto ldu = [&](const FrameCol& q, int k) {
        return (q.u && k >= 1 && k <= ns - 2) ? tile::ld(a.U0 + (size_t)(k - 1) * upl, q.ou)
                                              : 0.0;
    };
    auto ldv = [&](const FrameCol& q, int k) {
        return (q.v && k >= 1 && k <= ns - 2) ? tile::ld(a.V0 + (size_t)(k - 1) * vpl, q.ov)
                                              : 0.0;
    };
    auto ldw = [&](const FrameCol& q, int k) {
        return (q.w && k >= 0 && k <= ns - 2) ? tile::ld(a.W0 + (size_t)k * wpl, q.ow) : 0.0;
    };

    const int z0 = bz * a.zc, z1 = min(z0 + a.zc, ns);
    const int p0 = max(z0 - 1, 0), p1 = min(z1, ns - 1);
    // own column at the top of iteration p: c^n (p-1, p), c1 (p-2, p-1), U, V (p-1),
    // W (p-2, p-1), the W neighbours (L+1, L+RX) of plane p-2; n*: c^n at p+1 and the
    // faces at p, loaded one iteration ahead (first pass; later passes read the pass
    // input field as nf)
    double c0m = 0, c0c = 0, nc0 = 0, nf = 0, c1m = 0, c1c = 0, uc = 0, vc = 0;
    double wmm = 0, wm1 = ldw(P, p0 - 1), wmr = 0, wmu = 0;
    double nu = ldu(P, p0), nv = ldv(P, p0), nw = ldw(P, p0);
    double hc0m = 0, hc0c = 0, hnc0 = 0, hnf = 0, hwm1 = 0;
    double hnu = 0, hnv = 0, hnw = 0;
    if (DONOR) {
        c0m = ldc(P, p0 - 1);
        c0c = ldc(P, p0);
        nc0 = ldc(P, p0 + 1);
    } else {
        nf = ldc(P, p0);
    }
    if (hl) {
        if (DONOR) {
            hc0m = ldc(H, p0 - 1);
            hc0c = ldc(H, p0);
            hnc0 = ldc(H, p0 + 1);
        } else if (h1) {
            hnf = ldc(H, p0);
        }
        hnu = ldu(H, p0);
        hnv = ldv(H, p0);
        if (h1) {
            hwm1 = ldw(H, p0 - 1);
            hnw = ldw(H, p0);
        }
    }
    for (int pl = p0; pl <= p1; pl++) {
        const int p = __builtin_amdgcn_readfirstlane(pl);
        P.keep();
        H.keep();
        const int s3 = (p + 3) % 3;
        const double c0n = nc0, un = nu, vn = nv, wn = nw, fld = nf;
        const double hc0n = hnc0, hun = hnu, hvn = hnv, hwn = hnw, hfld = hnf;
        if (p < p1) {  // the next iteration's values: c^n at p + 2, faces at p + 1
            if (DONOR)
                nc0 = ldc(P, p + 2);
            else
                nf = ldc(P, p + 1);
            nu = ldu(P, p + 1);
            nv = ldv(P, p + 1);
            nw = ldw(P, p + 1);
            if (hl) {
                if (DONOR)
                    hnc0 = ldc(H, p + 2);
                else if (h1)
                    hnf = ldc(H, p + 1);
                hnu = ldu(H, p + 1);
                hnv = ldv(H, p + 1);
                if (h1) hnw = ldw(H, p + 1);
            }
        }
        if (DONOR) s_c0[L] = c0c;
        s_u[s3][L] = un;
        s_v[s3][L] = vn;
        if (hl) {
            if (DONOR) s_c0[HL] = hc0c;
            s_u[s3][HL] = hun;
            s_v[s3][HL] = hvn;
        }
        tile::lds_barrier();
        // donor field of plane p (smolar.c:51-53: frame cells keep c^n)
        double c1p;
        if (DONOR) {
            c1p = c0c;
            if (P.c && !frame_xy && p != 0 && p != ns - 1)
                c1p = donor(c0c, s_c0[L - 1], s_c0[L + 1], s_c0[L - RX], s_c0[L + RX], c0m,
                            c0n, s_u[s3][L - 1], un, s_v[s3][L - RX], vn, wm1, wn);
            if (P.c && p >= z0 && p < z1) tile::st(a.c1out + (size_t)p * np, P.oc, c1p);
        } else {
            c1p = fld;
        }
        s_c1[s3][L] = c1p;
        s_w[L] = wm1;
        if (h1) {
            double hc1 = DONOR ? hc0c : hfld;
            if (DONOR && H.c && !hframe_xy && p != 0 && p != ns - 1)
                hc1 = donor(hc0c, s_c0[HL - 1], s_c0[HL + 1], s_c0[HL - RX], s_c0[HL + RX],
                            hc0m, hc0n, s_u[s3][HL - 1], hun, s_v[s3][HL - RX], hvn, hwm1,
                            hwn);
            s_c1[s3][HL] = hc1;
            s_w[HL] = hwm1;
        }
        tile::lds_barrier();
        const int k = p - 1;  // plane whose faces are computed now
        const double wkr = s_w[L + 1], wku = s_w[L + RX];
        if (k >= z0 && k < z1 && k >= 0) {
            const int m3 = (p + 1) % 3, c3 = (p + 2) % 3;  // slots of planes k-1, k
            const double* cm = s_c1[m3];
            const double* cc = s_c1[c3];
            const double* cq = s_c1[s3];
            const double* uk = s_u[c3];
            const double* uq = s_u[s3];
            const double* vk = s_v[c3];
            const double* vq = s_v[s3];
            const double c = c1c;
            if (P.u && k >= 1 && k <= ns - 2) {  // x-face (uu.c:56-99)
                const double av = uc;
                double div, sv, sw;
                if (gx == 0 || gx == nx - 2) {  // reduced pattern at column ii (uu.c:92-97)
                    const int o = gx == 0 ? 1 : 0;  // ii - gx
                    div = 0.0;
                    sv = 0.5 * (vk[L + o - RX] + vk[L + o]);
                    sw = 0.5 * ((o ? wmr : wmm) + (o ? wkr : wm1));
                } else {
                    double vr = vk[L + 1 - RX], vru = vk[L + 1], vu = vc, vcc = vk[L - RX];
                    double wr = wmr, wrf = wkr, wf = wm1, wc = wmm;
                    sv = 0.25 * (vr + vru + vu + vcc);
                    sw = 0.25 * (wr + wrf + wf + wc);
                    div = 0.25 * av *
                          (uk[L + 1] - uk[L - 1] + vr - vru - vu + vcc + wr - wrf - wf + wc);
                }
                double cr = cc[L + 1], cru = cc[L + 1 + RX], cu = cc[L + RX];
                double cd = cc[L - RX], crd = cc[L + 1 - RX];
                double crf = cq[L + 1], cf = c1p, cb = c1m, crb = cm[L + 1];
                tile::st(a.U1 + (size_t)(k - 1) * upl, P.ou,
                         antidiff(div, av, cr, c, sv, cru + cu - cd - crd, cru + cu + cd + crd,
                                  sw, crf + cf - cb - crb, crf + cf + cb + crb));
            }
            const bool yedge = gy == 0 || gy == ny - 2;
            if (P.v && k >= 1 && k <= ns - 2 && !yedge && gx == 1) {
                const size_t o = (size_t)(k - 1) * vpl;
                tile::st(a.V1 + o, P.ov, a.q7 ? tile::ld(a.q7 + o, P.ov) : 0.0);
            } else if (P.v && k >= 1 && k <= ns - 2) {  // y-face (vv.c:58-104)
                const double av = vc;
                double div, su, sw;
                if (yedge) {  // reduced pattern at row jj (vv.c:97-103)
                    const int o = gy == 0 ? RX : 0;  // LDS offset of row jj
                    div = 0.0;
                    su = 0.5 * (uk[L - 1 + o] + (o ? uk[L + RX] : uc));
                    sw = 0.5 * ((o ? wmu : wmm) + (o ? wku : wm1));
                } else {
                    double uu = uk[L - 1 + RX], uru = uk[L + RX], ur = uc, ucc = uk[L - 1];
                    double ulu = uk[L - 2 + RX], ul = uk[L - 2];
                    double wu = wmu, wuf = wku, wf = wm1, wc = wmm;
                    su = 0.25 * (uu + uru + ur + ucc);
                    sw = 0.25 * (wu + wuf + wf + wc);
                    div = 0.25 * av *
                          (vk[L + RX] - vk[L - RX] + wu - wuf - wf + wc + uu - ulu - ul + ucc);
                }
                double cu = cc[L + RX], cru = cc[L + 1 + RX], cr = cc[L + 1];
                double cl = cc[L - 1], clu = cc[L - 1 + RX];
                double cuf = cq[L + RX], cf = c1p, cb = c1m, cub = cm[L + RX];
                tile::st(a.V1 + (size_t)(k - 1) * vpl, P.ov,
                         antidiff(div, av, cu, c, su, cru + cr - cl - clu, cru + cr + cl + clu,
                                  sw, cuf + cf - cb - cub, cuf + cf + cb + cub));
            }
            if (P.w && k <= ns - 2) {  // z-face (ww.c:222-270)
                const double av = wm1;
                double div, su, sv;
                if (k == 0 || k == ns - 2) {  // reduced pattern at plane kk (ww.c:101-110)
                    const bool up = k == 0;  // kk = k + 1
                    div = 0.0;
                    su = 0.5 * ((up ? uq : uk)[L - 1] + (up ? un : uc));
                    sv = 0.5 * ((up ? vq : vk)[L - RX] + (up ? vn : vc));
                } else {
                    double ur = uc, urf = un, uf = uq[L - 1], ucc = uk[L - 1];
                    double vf = vq[L - RX], vuf = vn, vu = vc, vcc = vk[L - RX];
                    su = 0.25 * (ur + urf + uf + ucc);
                    sv = 0.25 * (vf + vuf + vu + vcc);
                    div = 0.25 * av * (wn - wmm + uf - urf - ur + ucc + vf - vuf - vu + vcc);
                }
                double cf = c1p, crf = cq[L + 1], cr = cc[L + 1], cl = cc[L - 1];
                double clf = cq[L - 1], cuf = cq[L + RX], cu = cc[L + RX], cd = cc[L - RX];
                double cdf = cq[L - RX];
                tile::st(a.W1 + (size_t)k * wpl, P.ow,
                         antidiff(div, av, cf, c, su, crf + cr - cl - clf, crf + cr + cl + clf,
                                  sv, cuf + cu - cd - cdf, cuf + cu + cd + cdf));
            }
        }
        if (DONOR) {
            c0m = c0c;
            c0c = c0n;
            hc0m = hc0c;
            hc0c = hc0n;
        }
        c1m = c1c;
        c1c = c1p;
        uc = un;
        vc = vn;
        wmm = wm1;
        wm1 = wn;
        wmr = wkr;
        wmu = wku;
        hwm1 = hwn;
    }
}

// Launches k_frame_f1 over the whole grid: 32x8 tiles, z-chunks long enough to keep
// the 2-3 planes a chunk reads below itself a small share, short enough for ~2048
// workgroups.
int frame_f1(const mpd_grid_s* g, bool donor1, const double* cin, double* const prev[3],
             double* const out[3], double* c1out, const double* q7, hipStream_t s)
{
    constexpr int TX = 32, TY = 8;
    FrameF1Args a{};
    a.nx = g->nx;
    a.ny = g->ny;
    a.ns = g->nz;
    a.tiles_x = (a.nx + TX - 1) / TX;
    a.tiles_y = (a.ny + TY - 1) / TY;
    const long long cols = (long long)a.tiles_x * a.tiles_y;
    int zc = a.ns;
    while (zc > 8 && cols * ((a.ns + zc - 1) / zc) < 2048) zc = (zc + 1) / 2;
    a.zc = zc;
    const long long items = cols * ((a.ns + zc - 1) / zc);
    if (items > 0x7fffffff) return fail(MPD_ERR_UNSUPPORTED, "FRAME grid too large for one launch");
    if ((long long)a.nx * a.ny > (1LL << 29))  // 32-bit byte offsets within a plane
        return fail(MPD_ERR_UNSUPPORTED, "FRAME tiled path: planes of at most 2^29 cells");
    a.cin = cin;
    a.U0 = prev[0];
    a.V0 = prev[1];
    a.W0 = prev[2];
    a.c1out = c1out;
    a.U1 = out[0];
    a.V1 = out[1];
    a.W1 = out[2];
    a.q7 = q7;
    if (donor1)
        hipLaunchKernelGGL((k_frame_f1<TX, TY, true>), dim3((unsigned)items),
                           dim3(tile::Geo<TX, TY>::NT), 0, s, a);
    else
        hipLaunchKernelGGL((k_frame_f1<TX, TY, false>), dim3((unsigned)items),
                           dim3(tile::Geo<TX, TY>::NT), 0, s, a);
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
    // the tiled kernel forms the donor field together with the first pass's velocities;
    // mpd_set_path(g, 1) keeps the per-face kernels
    const bool tiled = g->path != 1 && g->iord > 1 && (long long)g->nx * g->ny <= (1LL << 29);
    double* set[2][3];
    for (int d = 0; d < 3; d++) set[0][d] = set[1][d] = g->face[d].p;
    if (g->iord > 1) {
        for (int d = 0; d < 3; d++) {
            if ((rc = g->fb[d].alloc(g->nface[d])) != MPD_OK) return rc;
            set[1][d] = g->fb[d].p;
            if (g->iord > 2) {  // set 0 is written from pass 2 on: a buffer of its own
                if ((rc = g->fa[d].alloc(g->nface[d])) != MPD_OK) return rc;
                set[0][d] = g->fa[d].p;
            }
            // The per-face kernels reproduce the reference's buffers: set 1 zero-filled,
            // set 0 a copy of the Courant numbers.  The tiled kernel writes every entry
            // of its output, the never-written column included, so needs neither.
            if (!tiled) {
                MPD_HIP(hipMemsetAsync(g->fb[d].p, 0, g->nface[d] * sizeof(double), s));
                if (g->iord > 2)
                    MPD_HIP(hipMemcpyAsync(g->fa[d].p, g->face[d].p,
                                           g->nface[d] * sizeof(double),
                                           hipMemcpyDeviceToDevice, s));
            }
        }
    }
    int cold = g->cur;
    for (int o = 0; o < g->iord; o++) {
        // pass o's velocities: set 1 on odd passes, set 0 on even ones (pass 0: the
        // Courant numbers or, with more than two passes on the per-face path, their copy)
        double** vel = set[o & 1];
        if (tiled) {
            if (o == 0) continue;  // formed in pass 1's kernel
            const bool first = o == 1;
            double* const courant[3] = {g->face[0].p, g->face[1].p, g->face[2].p};
            double* const* prev = first ? courant : set[(o - 1) & 1];
            if ((rc = frame_f1(g, first, g->c[cold].p, prev, vel,
                               first ? g->c[1 - cold].p : nullptr,
                               (o & 1) ? nullptr : g->face[1].p, s)))
                return rc;
            if (first) cold = 1 - cold;  // c1 is now the newest field
        } else if (o) {
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
