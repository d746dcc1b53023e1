// PERIODIC mode, staged path: one kernel per stage of the step, intermediates in HBM.
//
// Stages per step (iord passes):
//   donor            c1 = donor(c^n; Courant U,V,W)
//   per corrective pass p:
//     antidiff       U',V',W' from the pass input and the previous pass's velocities
//     beta           limiter ratios bup/bdn per cell (nonoscillatory option)
//     donor_lim      c_{p+1} = donor(c_p; limited U',V',W'), limited faces kept for
//                    the next pass; the last pass adds the emission increment.
// All arithmetic comes from mpd_formulas.hpp in the oracle's order, so this path is
// bit-identical to the oracle; it is the reference point for the fused path.
#include "mpd_formulas.hpp"
#include "mpd_grid.hpp"

namespace mpd {
using namespace exact;
namespace {

constexpr int TPB = 256;
inline unsigned nblocks(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// Periodic neighbour offsets of one cell.
struct Cell {
    size_t o, xl, xr, yd, yu, zb, zf;
    int i, j, k;
};

struct PerIdx {
    int nx, ny, nz;
    size_t np, n;
    __device__ Cell at(size_t t) const
    {
        Cell c;
        c.o = t;
        c.i = (int)(t % nx);
        c.j = (int)((t / nx) % ny);
        c.k = (int)(t / np);
        c.xl = t - c.i + (c.i == 0 ? nx - 1 : c.i - 1);
        c.xr = t - c.i + (c.i + 1 == nx ? 0 : c.i + 1);
        c.yd = t + ((size_t)(c.j == 0 ? ny - 1 : c.j - 1) - c.j) * nx;
        c.yu = t + ((size_t)(c.j + 1 == ny ? 0 : c.j + 1) - c.j) * nx;
        c.zb = t + ((size_t)(c.k == 0 ? nz - 1 : c.k - 1) - c.k) * np;
        c.zf = t + ((size_t)(c.k + 1 == nz ? 0 : c.k + 1) - c.k) * np;
        return c;
    }
    // offset of (i+di, j+dj, k+dk), wrapped, for |d| <= 2
    __device__ size_t off(const Cell& c, int di, int dj, int dk) const
    {
        int i = c.i + di, j = c.j + dj, k = c.k + dk;
        i += (i < 0) ? nx : 0;
        i -= (i >= nx) ? nx : 0;
        i += (i < 0) ? nx : 0;
        j += (j < 0) ? ny : 0;
        j -= (j >= ny) ? ny : 0;
        k += (k < 0) ? nz : 0;
        k -= (k >= nz) ? nz : 0;
        return (size_t)k * np + (size_t)j * nx + i;
    }
};

__global__ void k_per_donor(PerIdx x, const double* __restrict__ U, const double* __restrict__ V,
                            const double* __restrict__ W, const double* __restrict__ cin,
                            const double* __restrict__ q, double* __restrict__ cout)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= x.n) return;
    Cell c = x.at(t);
    double r = donor(cin[t], cin[c.xl], cin[c.xr], cin[c.yd], cin[c.yu], cin[c.zb], cin[c.zf],
                     U[c.xl], U[t], V[c.yd], V[t], W[c.zb], W[t]);
    if (q) r = r + q[t];
    cout[t] = r;
}

// The reference's interior antidiffusive formulas (uu.c:81-87, vv.c:85-92,
// ww.c:248-255) with wrapped indices; one thread writes the three faces above its cell.
__global__ void k_per_antidiff(PerIdx x, const double* __restrict__ c,
                               const double* __restrict__ U0, const double* __restrict__ V0,
                               const double* __restrict__ W0, double* __restrict__ U1,
                               double* __restrict__ V1, double* __restrict__ W1)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= x.n) return;
    Cell q = x.at(t);
    const size_t o = t;
    const size_t ip = q.xr, im = q.xl, jp = q.yu, jm = q.yd, kp = q.zf, km = q.zb;
    const size_t ipjm = x.off(q, 1, -1, 0), ipjp = x.off(q, 1, 1, 0);
    const size_t ipkm = x.off(q, 1, 0, -1), ipkp = x.off(q, 1, 0, 1);
    const size_t imjp = x.off(q, -1, 1, 0), imkp = x.off(q, -1, 0, 1);
    const size_t im2 = x.off(q, -2, 0, 0), im2jp = x.off(q, -2, 1, 0);
    const size_t jpkm = x.off(q, 0, 1, -1), jpkp = x.off(q, 0, 1, 1);
    const size_t jmkp = x.off(q, 0, -1, 1);
    {
        double a = U0[o];
        double vr = V0[ipjm], vru = V0[ip], vu = V0[o], vc = V0[jm];
        double wr = W0[ipkm], wrf = W0[ip], wf = W0[o], wc = W0[km];
        double sv = 0.25 * (vr + vru + vu + vc);
        double sw = 0.25 * (wr + wrf + wf + wc);
        double div = 0.25 * a * (U0[ip] - U0[im] + vr - vru - vu + vc + wr - wrf - wf + wc);
        double cc = c[o], cr = c[ip], cru = c[ipjp], cu = c[jp], cd = c[jm], crd = c[ipjm];
        double crf = c[ipkp], cf = c[kp], cb = c[km], crb = c[ipkm];
        U1[o] = antidiff(div, a, cr, cc, sv, cru + cu - cd - crd, cru + cu + cd + crd, sw,
                         crf + cf - cb - crb, crf + cf + cb + crb);
    }
    {
        double a = V0[o];
        double uu = U0[imjp], uru = U0[jp], ur = U0[o], uc = U0[im];
        double ulu = U0[im2jp], ul = U0[im2];
        double wu = W0[jpkm], wuf = W0[jp], wf = W0[o], wc = W0[km];
        double su = 0.25 * (uu + uru + ur + uc);
        double sw = 0.25 * (wu + wuf + wf + wc);
        double div = 0.25 * a * (V0[jp] - V0[jm] + wu - wuf - wf + wc + uu - ulu - ul + uc);
        double cc = c[o], cu = c[jp], cru = c[ipjp], cr = c[ip], cl = c[im], clu = c[imjp];
        double cuf = c[jpkp], cf = c[kp], cb = c[km], cub = c[jpkm];
        V1[o] = antidiff(div, a, cu, cc, su, cru + cr - cl - clu, cru + cr + cl + clu, sw,
                         cuf + cf - cb - cub, cuf + cf + cb + cub);
    }
    {
        double a = W0[o];
        double ur = U0[o], urf = U0[kp], uf = U0[imkp], uc = U0[im];
        double vf = V0[jmkp], vuf = V0[kp], vu = V0[o], vc = V0[jm];
        double su = 0.25 * (ur + urf + uf + uc);
        double sv = 0.25 * (vf + vuf + vu + vc);
        double div = 0.25 * a * (W0[kp] - W0[km] + uf - urf - ur + uc + vf - vuf - vu + vc);
        double cc = c[o], cf = c[kp], crf = c[ipkp], cr = c[ip], cl = c[im], clf = c[imkp];
        double cuf = c[jpkp], cu = c[jp], cd = c[jm], cdf = c[jmkp];
        W1[o] = antidiff(div, a, cf, cc, su, crf + cr - cl - clf, crf + cr + cl + clf, sv,
                         cuf + cu - cd - cdf, cuf + cu + cd + cdf);
    }
}

// Limiter ratios of one cell (Smolarkiewicz & Grabowski 1990; see oracle cpu_ref.c).
__global__ void k_per_beta(PerIdx x, const double* __restrict__ cn, const double* __restrict__ cp,
                           const double* __restrict__ U, const double* __restrict__ V,
                           const double* __restrict__ W, double* __restrict__ bup,
                           double* __restrict__ bdn)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= x.n) return;
    Cell q = x.at(t);
    const size_t nb[7] = {t, q.xl, q.xr, q.yd, q.yu, q.zb, q.zf};
    double hi = cn[t], lo = cn[t];
#pragma unroll
    for (int m = 0; m < 7; m++) {
        hi = mx2(hi, mx2(cn[nb[m]], cp[nb[m]]));
        lo = mn2(lo, mn2(cn[nb[m]], cp[nb[m]]));
    }
    double c = cp[t];
    double fxl = 0.5 * F2(cp[q.xl], c, U[q.xl]), fxr = 0.5 * F2(c, cp[q.xr], U[t]);
    double fyl = 0.5 * F2(cp[q.yd], c, V[q.yd]), fyr = 0.5 * F2(c, cp[q.yu], V[t]);
    double fzl = 0.5 * F2(cp[q.zb], c, W[q.zb]), fzr = 0.5 * F2(c, cp[q.zf], W[t]);
    beta(hi, lo, c, fxl, fxr, fyl, fyr, fzl, fzr, bup[t], bdn[t]);
}

// Donor pass with limited velocities computed on the fly from the unlimited faces and
// the ratios; optionally stores the limited upper faces (Ul) and adds q.
__global__ void k_per_donor_lim(PerIdx x, const double* __restrict__ U,
                                const double* __restrict__ V, const double* __restrict__ W,
                                const double* __restrict__ bup, const double* __restrict__ bdn,
                                const double* __restrict__ cin, const double* __restrict__ q,
                                double* __restrict__ cout, double* __restrict__ Ul,
                                double* __restrict__ Vl, double* __restrict__ Wl)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= x.n) return;
    Cell c = x.at(t);
    double bu = bup[t], bd = bdn[t];
    double uxr = limit_face(U[t], bu, bd, bup[c.xr], bdn[c.xr]);
    double uxl = limit_face(U[c.xl], bup[c.xl], bdn[c.xl], bu, bd);
    double vyr = limit_face(V[t], bu, bd, bup[c.yu], bdn[c.yu]);
    double vyl = limit_face(V[c.yd], bup[c.yd], bdn[c.yd], bu, bd);
    double wzr = limit_face(W[t], bu, bd, bup[c.zf], bdn[c.zf]);
    double wzl = limit_face(W[c.zb], bup[c.zb], bdn[c.zb], bu, bd);
    double r = donor(cin[t], cin[c.xl], cin[c.xr], cin[c.yd], cin[c.yu], cin[c.zb], cin[c.zf],
                     uxl, uxr, vyl, vyr, wzl, wzr);
    if (q) r = r + q[t];
    cout[t] = r;
    if (Ul) {
        Ul[t] = uxr;
        Vl[t] = vyr;
        Wl[t] = wzr;
    }
}

}  // namespace

int periodic_staged_step(mpd_grid_s* g)
{
    hipStream_t s = stream();
    int rc;
    const size_t n = g->ncell;
    PerIdx x{g->nx, g->ny, g->nz, (size_t)g->nx * g->ny, n};
    const unsigned nb = nblocks(n);
    const double* q = g->has_q ? g->q.p : nullptr;
    double* cin = g->c[g->cur].p;
    double* out = g->c[1 - g->cur].p;
    if (g->iord == 1) {
        hipLaunchKernelGGL(k_per_donor, dim3(nb), dim3(TPB), 0, s, x, g->face[0].p,
                           g->face[1].p, g->face[2].p, cin, q, out);
        MPD_HIP(hipGetLastError());
        g->cur = 1 - g->cur;
        return MPD_OK;
    }
    for (int d = 0; d < 3; d++) {
        if ((rc = g->fa[d].alloc(n)) != MPD_OK) return rc;
        if ((rc = g->fb[d].alloc(n)) != MPD_OK) return rc;
    }
    for (int b = 0; b < 2; b++)
        if ((rc = g->cp[b].alloc(n)) != MPD_OK) return rc;
    if (g->nonosc) {
        if ((rc = g->bup.alloc(n)) != MPD_OK) return rc;
        if ((rc = g->bdn.alloc(n)) != MPD_OK) return rc;
    }
    hipLaunchKernelGGL(k_per_donor, dim3(nb), dim3(TPB), 0, s, x, g->face[0].p, g->face[1].p,
                       g->face[2].p, cin, (const double*)nullptr, g->cp[0].p);
    const double* prev[3] = {g->face[0].p, g->face[1].p, g->face[2].p};
    const double* field = g->cp[0].p;
    for (int p = 1; p < g->iord; p++) {
        bool last = (p == g->iord - 1);
        double* dst = last ? out : g->cp[p & 1].p;
        // unlimited faces: with the limiter they always go to fa (the limited ones to
        // fb); without it fa/fb alternate so the previous pass's faces stay intact.
        mpd::DevBuf* tgt = (g->nonosc || (p & 1)) ? g->fa : g->fb;
        hipLaunchKernelGGL(k_per_antidiff, dim3(nb), dim3(TPB), 0, s, x, field, prev[0],
                           prev[1], prev[2], tgt[0].p, tgt[1].p, tgt[2].p);
        if (g->nonosc) {
            hipLaunchKernelGGL(k_per_beta, dim3(nb), dim3(TPB), 0, s, x, cin, field, tgt[0].p,
                               tgt[1].p, tgt[2].p, g->bup.p, g->bdn.p);
            double* keep[3] = {nullptr, nullptr, nullptr};
            if (!last)
                for (int d = 0; d < 3; d++) keep[d] = g->fb[d].p;
            hipLaunchKernelGGL(k_per_donor_lim, dim3(nb), dim3(TPB), 0, s, x, tgt[0].p,
                               tgt[1].p, tgt[2].p, g->bup.p, g->bdn.p, field,
                               last ? q : (const double*)nullptr, dst, keep[0], keep[1],
                               keep[2]);
            for (int d = 0; d < 3; d++) prev[d] = g->fb[d].p;
        } else {
            hipLaunchKernelGGL(k_per_donor, dim3(nb), dim3(TPB), 0, s, x, tgt[0].p, tgt[1].p,
                               tgt[2].p, field, last ? q : (const double*)nullptr, dst);
            for (int d = 0; d < 3; d++) prev[d] = tgt[d].p;
        }
        field = dst;
    }
    MPD_HIP(hipGetLastError());
    g->cur = 1 - g->cur;
    return MPD_OK;
}

}  // namespace mpd
