// PERIODIC mode, fused path: two LDS-tiled kernels per corrective pass.
//
//   F1  k_f1<DONOR>: [donor pass c1 = donor(c^n)] + antidiffusive velocities of c1
//       (the first corrective pass computes the donor field on the fly; later passes
//       read their input field).  Writes c1 (first pass) and the unlimited faces A.
//   F2  k_f2: limiter ratios + limited donor pass.  Writes the pass result (plus the
//       emission increment on the last pass) and, when another pass follows, the
//       limited faces it will use.
//
// Work decomposition.  A workgroup owns a TX x TY column tile of the (x, y) plane and a
// chunk of ZC planes; it marches up the chunk one plane per iteration (2.5-D streaming).
// Every thread owns one column of the tile plus a halo of 2 cells, loads that column's
// values plane by plane (coalesced along x), keeps the values it needs from the planes
// below and above in registers, and publishes the current plane in LDS so neighbours
// read in-plane values at fixed LDS offsets.  Threads of the tile interior are numbered
// first, so the halo threads form their own waves and sit out the expensive stages.
// Stage regions shrink inward: loads on the tile + 2 ring, the donor field / limiter
// ratios on the tile + 1 ring, the outputs on the tile.  Only the halo columns are
// computed twice (by two workgroups); in z nothing is recomputed except the two or
// three planes a chunk needs below its first output plane.
//
// Arithmetic comes from mpd_formulas.hpp in the oracle's order.  This file is compiled
// twice: MPD_ARITH=exact (-ffp-contract=off, IEEE division; bit-identical to the staged
// path and the oracle) and MPD_ARITH=fast (contracted multiply-adds, reciprocal-based
// division; <= 1e-12 normwise per step against the oracle).
#include "mpd_formulas.hpp"
#include "mpd_fused.hpp"
#include "mpd_grid.hpp"

#define MPD_CAT2(a, b) a##_##b
#define MPD_CAT(a, b) MPD_CAT2(a, b)

namespace mpd {
int ensure_fused_scratch(mpd_grid_s* g);
using namespace MPD_ARITH;
namespace {



template <int TX, int TY>
struct Geo {
    static constexpr int RX = TX + 4, RY = TY + 4, R = RX * RY;
    static constexpr int NT = ((R + 63) / 64) * 64;
    static constexpr int NTILE = TX * TY;
    // region cell of thread t: interior tile first, then the two halo rings
    __device__ static bool cell(int t, int& rx, int& ry)
    {
        if (t < NTILE) {
            rx = 2 + t % TX;
            ry = 2 + t / TX;
            return true;
        }
        int r = t - NTILE;
        if (r < 2 * RX) {
            ry = r / RX;
            rx = r % RX;
            return true;
        }
        r -= 2 * RX;
        if (r < 2 * RX) {
            ry = RY - 2 + r / RX;
            rx = r % RX;
            return true;
        }
        r -= 2 * RX;
        if (r < 4 * TY) {
            ry = 2 + r / 4;
            int s = r % 4;
            rx = (s < 2) ? s : RX - 4 + s;
            return true;
        }
        rx = ry = 0;
        return false;
    }
    __device__ static int ring(int rx, int ry)
    {
        int dx = rx < 2 ? 2 - rx : (rx > TX + 1 ? rx - TX - 1 : 0);
        int dy = ry < 2 ? 2 - ry : (ry > TY + 1 ? ry - TY - 1 : 0);
        return dx > dy ? dx : dy;
    }
};

// Workgroup barrier ordering LDS only.  A full __syncthreads() also fences global
// memory, which makes every wave wait for its outstanding global loads at the barrier
// and would defeat the one-plane-ahead prefetch below; these kernels exchange data
// between threads through LDS only.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ int wrapi(int a, int n)
{
    a %= n;
    return a < 0 ? a + n : a;
}

// tile (bx, by) and z-chunk bz of work item b of an instance
__device__ __forceinline__ void tile_of(const FusedArgs& a, int b, int& bx, int& by, int& bz)
{
    bx = b % a.tiles_x;
    by = (b / a.tiles_x) % a.tiles_y;
    bz = b / (a.tiles_x * a.tiles_y);
}

// Work items of a launch.  A single grid passes its descriptor by value (ninst = 0);
// a batch of instances passes a device table of descriptors and the first item of each
// instance.  Workgroups stride over the items (persistent when the grid is smaller than
// the item count).  When remap is set, items are dealt so that each of the 8 XCDs --
// which receive workgroups round robin -- works on one contiguous eighth of the item
// list, keeping halo-sharing neighbours in the same L2.
struct Work {
    const FusedArgs* table;
    const int* item0;
    int ninst, nitems, remap;
};

__device__ __forceinline__ int find_inst(const int* item0, int n, int item)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (item0[mid] <= item)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int item_of(const Work& w, int it)
{
    return w.remap ? (it % 8) * (w.nitems / 8) + it / 8 : it;
}

template <int TX, int TY, bool DONOR>
__device__ __forceinline__ void f1_tile(const FusedArgs& a, int item)
{
    using G = Geo<TX, TY>;
    constexpr int RX = G::RX, R = G::R;
    __shared__ double s_c0[2][R], s_c1[3][R], s_u[3][R], s_v[3][R], s_w[4][R];

    int rx, ry;
    const bool live = G::cell(threadIdx.x, rx, ry);
    const int ring = live ? G::ring(rx, ry) : 3;
    int bx, by, bz;
    tile_of(a, item, bx, by, bz);
    const int gx = wrapi(bx * TX + rx - 2, a.nx), gy = wrapi(by * TY + ry - 2, a.ny);
    const size_t col = (size_t)gy * a.nx + gx;
    const int L = ry * RX + rx;
    const bool store = ring == 0 && bx * TX + rx - 2 < a.nx && by * TY + ry - 2 < a.ny;
    const int z0 = bz * a.zc;
    const int z1 = min(z0 + a.zc, a.nz);
    const size_t np = (size_t)a.np;
    auto at = [&](int z) { return (size_t)wrapi(z, a.nz) * np + col; };

    // registers: own column of c^n (planes p-1, p) and W (plane p-1); the values an
    // iteration consumes are loaded one iteration ahead (n*) so that their latency
    // overlaps the previous plane's arithmetic
    double c0m = 0.0, c0c = 0.0, wm1 = 0.0;
    double nc0 = 0.0, nu = 0.0, nv = 0.0, nw = 0.0, nf = 0.0;
    auto load_next = [&](int p) {
        if (DONOR)
            nc0 = a.cin[at(p + 1)];
        else
            nf = a.cin[at(p)];
        nu = a.U[at(p)];
        nv = a.V[at(p)];
        nw = a.W[at(p)];
    };
    if (live) {
        if (DONOR) {
            c0m = a.cin[at(z0 - 2)];
            c0c = a.cin[at(z0 - 1)];
        }
        wm1 = a.W[at(z0 - 2)];
        load_next(z0 - 1);
    }
    for (int p = z0 - 1; p <= z1; p++) {
        const int s2 = (p + 3) % 2, s3 = (p + 3) % 3, s4 = (p + 4) % 4;
        const double c0n = nc0, un = nu, vn = nv, wn = nw, fld = nf;
        if (live) {
            if (p < z1) load_next(p + 1);
            if (DONOR) s_c0[s2][L] = c0c;
            s_u[s3][L] = un;
            s_v[s3][L] = vn;
            s_w[s4][L] = wn;
        }
        lds_barrier();
        if (DONOR) {
            if (ring <= 1) {
                const double* c0p = s_c0[s2];
                double c1 = donor(c0c, c0p[L - 1], c0p[L + 1], c0p[L - RX], c0p[L + RX], c0m,
                                  c0n, s_u[s3][L - 1], un, s_v[s3][L - RX], vn, wm1, wn);
                s_c1[s3][L] = c1;
                if (store && p >= z0 && p < z1) a.c1out[at(p)] = c1;
            }
        } else if (live) {
            s_c1[s3][L] = fld;
        }
        lds_barrier();
        const int k = p - 1;  // plane whose faces are computed now
        if (ring == 0 && k >= z0 && k < z1) {
            const int m3 = (p + 1) % 3, c3 = (p + 2) % 3;  // slots of planes k-1, k
            const double* cm = s_c1[m3];
            const double* cc = s_c1[c3];
            const double* cq = s_c1[s3];
            const double* uk = s_u[c3];
            const double* uq = s_u[s3];
            const double* vk = s_v[c3];
            const double* vq = s_v[s3];
            const double* wm = s_w[(p + 2) % 4];
            const double* wk = s_w[(p + 3) % 4];
            const double* wq = s_w[s4];
            const size_t o = at(k);
            {  // x-face (uu.c:81-87)
                double av = uk[L];
                double vr = vk[L + 1 - RX], vru = vk[L + 1], vu = vk[L], vc = vk[L - RX];
                double wr = wm[L + 1], wrf = wk[L + 1], wf = wk[L], wc = wm[L];
                double sv = 0.25 * (vr + vru + vu + vc);
                double sw = 0.25 * (wr + wrf + wf + wc);
                double div = 0.25 * av * (uk[L + 1] - uk[L - 1] + vr - vru - vu + vc + wr - wrf -
                                          wf + wc);
                double c = cc[L], cr = cc[L + 1], cru = cc[L + 1 + RX], cu = cc[L + RX];
                double cd = cc[L - RX], crd = cc[L + 1 - RX];
                double crf = cq[L + 1], cf = cq[L], cb = cm[L], crb = cm[L + 1];
                a.Ax[o] = antidiff(div, av, cr, c, sv, cru + cu - cd - crd, cru + cu + cd + crd,
                                   sw, crf + cf - cb - crb, crf + cf + cb + crb);
            }
            {  // y-face (vv.c:85-92)
                double av = vk[L];
                double uu = uk[L - 1 + RX], uru = uk[L + RX], ur = uk[L], uc = uk[L - 1];
                double ulu = uk[L - 2 + RX], ul = uk[L - 2];
                double wu = wm[L + RX], wuf = wk[L + RX], wf = wk[L], wc = wm[L];
                double su = 0.25 * (uu + uru + ur + uc);
                double sw = 0.25 * (wu + wuf + wf + wc);
                double div = 0.25 * av * (vk[L + RX] - vk[L - RX] + wu - wuf - wf + wc + uu - ulu -
                                          ul + uc);
                double c = cc[L], cu = cc[L + RX], cru = cc[L + 1 + RX], cr = cc[L + 1];
                double cl = cc[L - 1], clu = cc[L - 1 + RX];
                double cuf = cq[L + RX], cf = cq[L], cb = cm[L], cub = cm[L + RX];
                a.Ay[o] = antidiff(div, av, cu, c, su, cru + cr - cl - clu, cru + cr + cl + clu,
                                   sw, cuf + cf - cb - cub, cuf + cf + cb + cub);
            }
            {  // z-face (ww.c:248-255)
                double av = wk[L];
                double ur = uk[L], urf = uq[L], uf = uq[L - 1], uc = uk[L - 1];
                double vf = vq[L - RX], vuf = vq[L], vu = vk[L], vc = vk[L - RX];
                double su = 0.25 * (ur + urf + uf + uc);
                double sv = 0.25 * (vf + vuf + vu + vc);
                double div = 0.25 * av * (wq[L] - wm[L] + uf - urf - ur + uc + vf - vuf - vu + vc);
                double c = cc[L], cf = cq[L], crf = cq[L + 1], cr = cc[L + 1], cl = cc[L - 1];
                double clf = cq[L - 1], cuf = cq[L + RX], cu = cc[L + RX], cd = cc[L - RX];
                double cdf = cq[L - RX];
                a.Az[o] = antidiff(div, av, cf, c, su, crf + cr - cl - clf, crf + cr + cl + clf,
                                   sv, cuf + cu - cd - cdf, cuf + cu + cd + cdf);
            }
        }
        if (DONOR) {
            c0m = c0c;
            c0c = c0n;
        }
        wm1 = wn;
    }
}

template <int TX, int TY>
__device__ __forceinline__ void f2_tile(const FusedArgs& a, int item)
{
    using G = Geo<TX, TY>;
    constexpr int RX = G::RX, R = G::R;
    __shared__ double s_cn[2][R], s_cp[3][R], s_ax[3][R], s_ay[3][R], s_bu[2][R], s_bd[2][R];

    int rx, ry;
    const bool live = G::cell(threadIdx.x, rx, ry);
    const int ring = live ? G::ring(rx, ry) : 3;
    int bx, by, bz;
    tile_of(a, item, bx, by, bz);
    const int gx = wrapi(bx * TX + rx - 2, a.nx), gy = wrapi(by * TY + ry - 2, a.ny);
    const size_t col = (size_t)gy * a.nx + gx;
    const int L = ry * RX + rx;
    const bool store = ring == 0 && bx * TX + rx - 2 < a.nx && by * TY + ry - 2 < a.ny;
    const int z0 = bz * a.zc;
    const int z1 = min(z0 + a.zc, a.nz);
    const size_t np = (size_t)a.np;
    auto at = [&](int z) { return (size_t)wrapi(z, a.nz) * np + col; };

    // own-column registers: cn p-1,p ; cp p-2,p-1,p ; az p-2,p-1 ; beta p-2,p-1; the
    // values an iteration consumes are loaded one iteration ahead (n*)
    double cnm = 0, cnc = 0, cpm2 = 0, cpm = 0, cpc = 0, azm2 = 0, azm = 0;
    double bum2 = 0, bdm2 = 0, bum = 0, bdm = 0;
    double ncn = 0, ncp = 0, nax = 0, nay = 0, naz = 0;
    auto load_next = [&](int p) {
        ncn = a.cn[at(p + 1)];
        ncp = a.cp[at(p + 1)];
        nax = a.Ax[at(p)];
        nay = a.Ay[at(p)];
        naz = a.Az[at(p)];
    };
    if (live) {
        cnm = a.cn[at(z0 - 2)];
        cnc = a.cn[at(z0 - 1)];
        cpm = a.cp[at(z0 - 2)];
        cpc = a.cp[at(z0 - 1)];
        azm = a.Az[at(z0 - 2)];
        load_next(z0 - 1);
    }
    for (int p = z0 - 1; p <= z1; p++) {
        const int s2 = (p + 2) % 2, s3 = (p + 3) % 3;
        const double cnn = ncn, cpn = ncp, axn = nax, ayn = nay, azc = naz;
        if (live) {
            if (p < z1) load_next(p + 1);
            s_cn[s2][L] = cnc;
            s_cp[s3][L] = cpc;
            s_ax[s3][L] = axn;
            s_ay[s3][L] = ayn;
        }
        lds_barrier();
        double buc = 0, bdc = 0;
        if (ring <= 1) {  // limiter ratios of plane p (see cpu_ref.c)
            const double* cnp = s_cn[s2];
            const double* cpp = s_cp[s3];
            double hi = cnc, lo = cnc;
            double nbn[7] = {cnc, cnp[L - 1], cnp[L + 1], cnp[L - RX], cnp[L + RX], cnm, cnn};
            double nbp[7] = {cpc, cpp[L - 1], cpp[L + 1], cpp[L - RX], cpp[L + RX], cpm, cpn};
#pragma unroll
            for (int m = 0; m < 7; m++) {
                hi = mx2(hi, mx2(nbn[m], nbp[m]));
                lo = mn2(lo, mn2(nbn[m], nbp[m]));
            }
            const double* axp = s_ax[s3];
            const double* ayp = s_ay[s3];
            double c = cpc;
            double fxl = 0.5 * F2(nbp[1], c, axp[L - 1]), fxr = 0.5 * F2(c, nbp[2], axn);
            double fyl = 0.5 * F2(nbp[3], c, ayp[L - RX]), fyr = 0.5 * F2(c, nbp[4], ayn);
            double fzl = 0.5 * F2(cpm, c, azm), fzr = 0.5 * F2(c, cpn, azc);
            beta(hi, lo, c, fxl, fxr, fyl, fyr, fzl, fzr, buc, bdc);
            s_bu[s2][L] = buc;
            s_bd[s2][L] = bdc;
        }
        lds_barrier();
        const int k = p - 1;
        if (ring == 0 && k >= z0 && k < z1) {
            const int k2 = (p + 1) % 2, k3 = (p + 2) % 3;  // slots of plane k
            const double* bu = s_bu[k2];
            const double* bd = s_bd[k2];
            const double* ax = s_ax[k3];
            const double* ay = s_ay[k3];
            const double* cq = s_cp[k3];
            double b0u = bu[L], b0d = bd[L];
            double uxr = limit_face(ax[L], b0u, b0d, bu[L + 1], bd[L + 1]);
            double uxl = limit_face(ax[L - 1], bu[L - 1], bd[L - 1], b0u, b0d);
            double vyr = limit_face(ay[L], b0u, b0d, bu[L + RX], bd[L + RX]);
            double vyl = limit_face(ay[L - RX], bu[L - RX], bd[L - RX], b0u, b0d);
            double wzr = limit_face(azm, b0u, b0d, buc, bdc);
            double wzl = limit_face(azm2, bum2, bdm2, b0u, b0d);
            double r = donor(cq[L], cq[L - 1], cq[L + 1], cq[L - RX], cq[L + RX], cpm2, cpc, uxl,
                             uxr, vyl, vyr, wzl, wzr);
            if (store) {
                const size_t o = at(k);
                if (a.q) r = r + a.q[o];
                a.cout[o] = r;
                if (a.Lx) {
                    a.Lx[o] = uxr;
                    a.Ly[o] = vyr;
                    a.Lz[o] = wzr;
                }
            }
        }
        cnm = cnc;
        cnc = cnn;
        cpm2 = cpm;
        cpm = cpc;
        cpc = cpn;
        azm2 = azm;
        azm = azc;
        bum2 = bum;
        bdm2 = bdm;
        bum = buc;
        bdm = bdc;
    }
}

template <int TX, int TY, bool DONOR>
__global__ void __launch_bounds__((Geo<TX, TY>::NT)) k_f1(FusedArgs a, Work w)
{
    for (int it = blockIdx.x; it < w.nitems; it += gridDim.x) {
        int item = item_of(w, it);
        FusedArgs x = a;
        if (w.ninst) {
            int i = find_inst(w.item0, w.ninst, item);
            x = w.table[i];
            item -= w.item0[i];
        }
        f1_tile<TX, TY, DONOR>(x, item);
        lds_barrier();  // the next item reuses the LDS planes
    }
}

template <int TX, int TY>
__global__ void __launch_bounds__((Geo<TX, TY>::NT)) k_f2(FusedArgs a, Work w)
{
    for (int it = blockIdx.x; it < w.nitems; it += gridDim.x) {
        int item = item_of(w, it);
        FusedArgs x = a;
        if (w.ninst) {
            int i = find_inst(w.item0, w.ninst, item);
            x = w.table[i];
            item -= w.item0[i];
        }
        f2_tile<TX, TY>(x, item);
        lds_barrier();
    }
}

// The staged donor (mpd_periodic.hip) serves the unlimited corrective pass.
constexpr int TPB = 256;
__global__ void k_donor_faces(int nx, int ny, int nz, const double* __restrict__ U,
                              const double* __restrict__ V, const double* __restrict__ W,
                              const double* __restrict__ cin, const double* __restrict__ q,
                              double* __restrict__ cout)
{
    const size_t np = (size_t)nx * ny, n = np * nz;
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= n) return;
    int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / np);
    size_t xl = t - i + (i == 0 ? nx - 1 : i - 1), xr = t - i + (i + 1 == nx ? 0 : i + 1);
    size_t yd = t + ((size_t)(j == 0 ? ny - 1 : j - 1) - j) * nx;
    size_t yu = t + ((size_t)(j + 1 == ny ? 0 : j + 1) - j) * nx;
    size_t zb = t + ((size_t)(k == 0 ? nz - 1 : k - 1) - k) * np;
    size_t zf = t + ((size_t)(k + 1 == nz ? 0 : k + 1) - k) * np;
    double r = donor(cin[t], cin[xl], cin[xr], cin[yd], cin[yu], cin[zb], cin[zf], U[xl], U[t],
                     V[yd], V[t], W[zb], W[t]);
    if (q) r = r + q[t];
    cout[t] = r;
}

// ---------------------------------------------------------------- dispatch ---

template <int TX, int TY>
int launch_pair(const FusedArgs& a, const Work& w, int grid, bool f1, bool donor, hipStream_t s)
{
    using G = Geo<TX, TY>;
    if (f1) {
        if (donor)
            hipLaunchKernelGGL((k_f1<TX, TY, true>), dim3(grid), dim3(G::NT), 0, s, a, w);
        else
            hipLaunchKernelGGL((k_f1<TX, TY, false>), dim3(grid), dim3(G::NT), 0, s, a, w);
    } else {
        hipLaunchKernelGGL((k_f2<TX, TY>), dim3(grid), dim3(G::NT), 0, s, a, w);
    }
    MPD_HIP(hipGetLastError());
    return MPD_OK;
}

int launch_any(const FusedArgs& a, const Work& w, int grid, int tx, int ty, bool f1, bool donor,
               hipStream_t s)
{
#define MPD_TILE(X, Y) \
    if (tx == X && ty == Y) return launch_pair<X, Y>(a, w, grid, f1, donor, s);
    MPD_FUSED_TILES(MPD_TILE)
#undef MPD_TILE
    return fail(MPD_ERR_ARG, "tile %dx%d is not one of the compiled tile shapes", tx, ty);
}

// single grid: one work item per workgroup
int launch(FusedArgs& a, int tx, int ty, bool f1, bool donor, hipStream_t s)
{
    fused_geometry(a, tx, ty);
    const int nb = a.tiles_x * a.tiles_y * a.nchunks;
    Work w{nullptr, nullptr, 0, nb, nb % 8 == 0 ? 1 : 0};
    return launch_any(a, w, nb, tx, ty, f1, donor, s);
}

// z-chunk length: enough workgroups to fill 256 CUs several times over, long enough
// that the 2-3 extra planes below each chunk stay a small fraction.
int choose_zc(const mpd_grid_s* g, int tx, int ty)
{
    long long cols = (long long)((g->nx + tx - 1) / tx) * ((g->ny + ty - 1) / ty);
    int zc = g->nz;
    while (zc > 8 && cols * ((g->nz + zc - 1) / zc) < 2048) zc = (zc + 1) / 2;
    return zc;
}

}  // namespace

int MPD_CAT(fused_batch_launch, MPD_ARITH)(int tx, int ty, bool f1, bool donor,
                                           const FusedArgs* table, const int* item0, int ninst,
                                           int nitems, int grid, hipStream_t s)
{
    Work w{table, item0, ninst, nitems, (nitems % 8 == 0 && grid % 8 == 0) ? 1 : 0};
    FusedArgs none{};
    return launch_any(none, w, grid, tx, ty, f1, donor, s);
}

int MPD_CAT(fused_resident_blocks, MPD_ARITH)(int tx, int ty, bool f1)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
#define MPD_TILE(X, Y)                                                                      \
    if (tx == X && ty == Y)                                                                 \
        e = f1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_f1<X, Y, true>,         \
                                                             Geo<X, Y>::NT, 0)              \
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_f2<X, Y>, Geo<X, Y>::NT, 0);
    MPD_FUSED_TILES(MPD_TILE)
#undef MPD_TILE
    return e == hipSuccess ? n : 0;
}

int MPD_CAT(periodic_fused_step, MPD_ARITH)(mpd_grid_s* g)
{
    hipStream_t s = stream();
    int rc;
    const size_t n = g->ncell;
    if ((rc = ensure_fused_scratch(g))) return rc;
    int tx = g->tile_x ? g->tile_x : 32, ty = g->tile_y ? g->tile_y : 8;
    FusedArgs a{};
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.np = (long long)g->nx * g->ny;
    a.zc = g->zc ? g->zc : choose_zc(g, tx, ty);
    const double* q = g->has_q ? g->q.p : nullptr;
    double* cin = g->c[g->cur].p;
    double* out = g->c[1 - g->cur].p;
    const double* prev[3] = {g->face[0].p, g->face[1].p, g->face[2].p};
    const double* field = cin;
    for (int p = 1; p < g->iord; p++) {
        const bool last = (p == g->iord - 1);
        double* dst = last ? out : g->cp[p & 1].p;
        mpd::DevBuf* tgt = (g->nonosc || (p & 1)) ? g->fa : g->fb;
        // F1: the first pass forms the donor field in the same kernel
        a.cin = field;
        a.U = prev[0];
        a.V = prev[1];
        a.W = prev[2];
        a.c1out = g->cp[0].p;
        a.Ax = tgt[0].p;
        a.Ay = tgt[1].p;
        a.Az = tgt[2].p;
        if ((rc = launch(a, tx, ty, true, p == 1, s))) return rc;
        trace_mark(g);
        const double* pass_in = (p == 1) ? g->cp[0].p : field;
        if (g->nonosc) {
            a.cn = cin;
            a.cp = pass_in;
            a.q = last ? q : nullptr;
            a.cout = dst;
            a.Lx = last ? nullptr : g->fb[0].p;
            a.Ly = last ? nullptr : g->fb[1].p;
            a.Lz = last ? nullptr : g->fb[2].p;
            if ((rc = launch(a, tx, ty, false, false, s))) return rc;
            trace_mark(g);
            for (int d = 0; d < 3; d++) prev[d] = g->fb[d].p;
        } else {
            hipLaunchKernelGGL(k_donor_faces, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0,
                               s, g->nx, g->ny, g->nz, tgt[0].p, tgt[1].p, tgt[2].p, pass_in,
                               last ? q : (const double*)nullptr, dst);
            MPD_HIP(hipGetLastError());
            trace_mark(g);
            for (int d = 0; d < 3; d++) prev[d] = tgt[d].p;
        }
        field = dst;
    }
    g->cur = 1 - g->cur;
    return MPD_OK;
}

}  // namespace mpd
