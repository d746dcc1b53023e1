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
// Each thread owns one column of the tile and at most one column of the 2-cell halo
// around it.  It loads both columns plane by plane (coalesced along x), keeps the values
// it needs from the planes below and above in registers, and publishes the current plane
// in LDS so neighbours read in-plane values at fixed LDS offsets.  Every wave computes
// the tile's expensive stages (no wave holds only halo cells); the threads whose halo
// cell lies on ring 1 -- the lowest-numbered ones -- also compute the donor field or
// the limiter ratios there.  Only the ring-1 cells are computed twice (by two
// workgroups); in z nothing is recomputed except the two or three planes a chunk needs
// below its first output plane.
//
// Arithmetic comes from mpd_formulas.hpp in the oracle's order.  This file is compiled
// twice: MPD_ARITH=exact (-ffp-contract=off, IEEE division; bit-identical to the staged
// path and the oracle) and MPD_ARITH=fast (contracted multiply-adds, reciprocal-based
// division; <= 1e-12 normwise per step against the oracle).
#include <hip/hip_ext.h>

#include <algorithm>
#include <type_traits>

#include "mpd_formulas.hpp"
#include "mpd_fused.hpp"
#include "mpd_grid.hpp"
#include "mpd_tile.hpp"

#define MPD_CAT2(a, b) a##_##b
#define MPD_CAT(a, b) MPD_CAT2(a, b)

namespace mpd {
int ensure_fused_scratch(mpd_grid_s* g);
using namespace MPD_ARITH;
namespace {



using namespace tile;

// A store of the fused kernels: streaming in the NT instantiations (chosen at launch from
// FusedArgs::nt; a run-time branch on the flag measured no gain at all).
template <bool NT>
__device__ __forceinline__ void put(double* plane, unsigned colb, double v)
{
    if constexpr (NT)
        st_nt(plane, colb, v);
    else
        st(plane, colb, v);
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

// The descriptor table and the item offsets are read through the constant address
// space: the reads are scalar loads and their results wave-uniform.  (Through a generic
// pointer the compiler must assume each lane may read a different value, and every
// quantity derived from the descriptor -- the plane loop included -- would be computed
// per lane.)
typedef const __attribute__((address_space(4))) int ConstInt;
typedef const __attribute__((address_space(4))) FusedArgs ConstArgs;

__device__ __forceinline__ int find_inst(ConstInt* item0, int n, int item)
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

// F1 over one work item.  Iteration p: publish c^n(p), U(p), V(p) -> barrier -> donor
// field c1(p) on the tile and ring 1, publish it with W(p-1) -> barrier -> antidiffusive
// velocities of the faces above/right of plane k = p - 1 on the tile.
// LDS slots: c^n and W one plane each (each is read only between the two barriers that
// follow its write), c1, U and V three planes each (plane p is written while slower waves
// may still read planes p-2 and p-1).  The own-column values of the planes below and
// above, and the W neighbours of plane k-1, are carried in registers.
template <int TX, int TY, bool DONOR, bool NT>
__device__ __forceinline__ void f1_tile(const FusedArgs& a, int item)
{
    using G = Geo<TX, TY>;
    constexpr int RX = G::RX, R = G::R;
    __shared__ double s_c0[R], s_c1[3][R], s_u[3][R], s_v[3][R], s_w[R];

    const int t = G::role(item);
    int bx, by, bz;
    tile_of(a, item, bx, by, bz);
    // own tile cell
    const int rx = 2 + t % TX, ry = 2 + t / TX;
    const int L = ry * RX + rx;
    unsigned colb = column<TX, TY>(a.nx, a.ny, bx, by, rx, ry);
    const bool store = bx * TX + rx - 2 < a.nx && by * TY + ry - 2 < a.ny;
    // own halo cell: ring 1 (loads, donor) or ring 2 (loads), if any
    const bool hl = t < G::NH, h1 = t < G::H1;
    int hx = 0, hy = 0;
    if (hl) G::halo(t, hx, hy);
    const int HL = hy * RX + hx;
    unsigned hcolb = column<TX, TY>(a.nx, a.ny, bx, by, hx, hy);

    const int z0 = bz * a.zc;
    const int z1 = min(z0 + a.zc, a.nz);
    // planes p + 2, p + 1, p, p - 1 of iteration p
    ZCursor zp2(z0 + 1, a.nz, a.np), zp1(z0, a.nz, a.np), zp0(z0 - 1, a.nz, a.np), zm1(z0 - 2, a.nz, a.np);

    // tile cell registers: c^n (p-1, p), c1 (p-2, p-1), U, V (p-1), W (p-2, p-1), the W
    // neighbours (L+1, L+RX) of plane p-2; n*: the next iteration's values, loaded one
    // iteration ahead so that their latency overlaps this plane's arithmetic
    double c0m = 0, c0c = 0, c1m = 0, c1c = 0, uc = 0, vc = 0, wmm = 0, wm1 = 0, wmr = 0, wmu = 0;
    double nc0 = 0, nu = 0, nv = 0, nw = 0, nf = 0;
    // halo cell registers
    double hc0m = 0, hc0c = 0, hwm1 = 0;
    double hnc0 = 0, hnu = 0, hnv = 0, hnw = 0, hnf = 0;
    if (DONOR) {
        c0m = ld(a.cin + zm1.off, colb);
        c0c = ld(a.cin + zp0.off, colb);
        nc0 = ld(a.cin + zp1.off, colb);
    } else {
        nf = ld(a.cin + zp0.off, colb);
    }
    wm1 = ld(a.W + zm1.off, colb);
    nu = ld(a.U + zp0.off, colb);
    nv = ld(a.V + zp0.off, colb);
    nw = ld(a.W + zp0.off, colb);
    if (hl) {
        if (DONOR) {
            hc0m = ld(a.cin + zm1.off, hcolb);
            hc0c = ld(a.cin + zp0.off, hcolb);
            hnc0 = ld(a.cin + zp1.off, hcolb);
        } else if (h1) {
            hnf = ld(a.cin + zp0.off, hcolb);
        }
        hnu = ld(a.U + zp0.off, hcolb);
        hnv = ld(a.V + zp0.off, hcolb);
        if (h1) {
            hwm1 = ld(a.W + zm1.off, hcolb);
            hnw = ld(a.W + zp0.off, hcolb);
        }
    }
    for (int pl = z0 - 1; pl <= z1; pl++) {
        // the plane index is the same in every lane; without the hint the compiler keeps
        // it, and the LDS slots derived from it, in vector registers
        const int p = __builtin_amdgcn_readfirstlane(pl);
        keep32(colb);
        keep32(hcolb);
        const int s3 = (p + 3) % 3;
        // raised issue priority from the next plane's loads to the first barrier, so that
        // a wave's loads and LDS writes go out ahead of other waves' arithmetic (measured
        // 0.3 % on the step; the same in f2_tile 0.5 %)
        __builtin_amdgcn_s_setprio(2);
        const double c0n = nc0, un = nu, vn = nv, wn = nw, fld = nf;
        const double hc0n = hnc0, hun = hnu, hvn = hnv, hwn = hnw, hfld = hnf;
        if (p < z1) {  // the next iteration's values: c^n at p + 2, faces at p + 1
            if (DONOR)
                nc0 = ld(a.cin + zp2.off, colb);
            else
                nf = ld(a.cin + zp1.off, colb);
            nu = ld(a.U + zp1.off, colb);
            nv = ld(a.V + zp1.off, colb);
            nw = ld(a.W + zp1.off, colb);
            if (hl) {
                if (DONOR)
                    hnc0 = ld(a.cin + zp2.off, hcolb);
                else if (h1)
                    hnf = ld(a.cin + zp1.off, hcolb);
                hnu = ld(a.U + zp1.off, hcolb);
                hnv = ld(a.V + zp1.off, hcolb);
                if (h1) hnw = ld(a.W + zp1.off, hcolb);
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
        lds_barrier();
        __builtin_amdgcn_s_setprio(0);
        double c1p;
        if (DONOR) {
            c1p = donor(c0c, s_c0[L - 1], s_c0[L + 1], s_c0[L - RX], s_c0[L + RX], c0m, c0n,
                        s_u[s3][L - 1], un, s_v[s3][L - RX], vn, wm1, wn);
            if (store && p >= z0 && p < z1) put<NT>(a.c1out + zp0.off, colb, c1p);
        } else {
            c1p = fld;
        }
        s_c1[s3][L] = c1p;
        s_w[L] = wm1;
        if (h1) {
            const double hc1 =
                DONOR ? donor(hc0c, s_c0[HL - 1], s_c0[HL + 1], s_c0[HL - RX], s_c0[HL + RX],
                              hc0m, hc0n, s_u[s3][HL - 1], hun, s_v[s3][HL - RX], hvn, hwm1, hwn)
                      : hfld;
            s_c1[s3][HL] = hc1;
            s_w[HL] = hwm1;
        }
        lds_barrier();
        const int k = p - 1;  // plane whose faces are computed now
        const double wkr = s_w[L + 1], wku = s_w[L + RX];
        if (k >= z0 && k < z1) {
            const int m3 = (p + 1) % 3, c3 = (p + 2) % 3;  // slots of planes k-1, k
            const double* cm = s_c1[m3];
            const double* cc = s_c1[c3];
            const double* cq = s_c1[s3];
            const double* uk = s_u[c3];
            const double* uq = s_u[s3];
            const double* vk = s_v[c3];
            const double* vq = s_v[s3];
            const double c = c1c;
            {  // x-face (uu.c:81-87)
                double av = uc;
                double vr = vk[L + 1 - RX], vru = vk[L + 1], vu = vc, vcc = vk[L - RX];
                double wr = wmr, wrf = wkr, wf = wm1, wc = wmm;
                double cr = cc[L + 1], cru = cc[L + 1 + RX], cu = cc[L + RX];
                double cd = cc[L - RX], crd = cc[L + 1 - RX];
                double crf = cq[L + 1], cf = c1p, cb = c1m, crb = cm[L + 1];
#ifdef MPD_FAST_DIV
                // fast arithmetic: stroke pairs and cell pairs shared between the averages,
                // the divergence term and the ratio terms (the same sums, reassociated)
                const double v1 = vr + vcc, v2 = vru + vu, w1 = wr + wc, w2 = wrf + wf;
                const double b1 = cru + cu, b2 = cd + crd, e1 = crf + cf, e2 = cb + crb;
                double sv = 0.25 * (v1 + v2);
                double sw = 0.25 * (w1 + w2);
                double div = 0.25 * av * ((uk[L + 1] - uk[L - 1]) + (v1 - v2) + (w1 - w2));
                put<NT>(a.Ax + zm1.off, colb,
                   antidiff(div, av, cr, c, sv, b1 - b2, b1 + b2, sw, e1 - e2, e1 + e2));
#else
                double sv = 0.25 * (vr + vru + vu + vcc);
                double sw = 0.25 * (wr + wrf + wf + wc);
                double div = 0.25 * av * (uk[L + 1] - uk[L - 1] + vr - vru - vu + vcc + wr - wrf -
                                          wf + wc);
                put<NT>(a.Ax + zm1.off, colb,
                   antidiff(div, av, cr, c, sv, cru + cu - cd - crd, cru + cu + cd + crd, sw,
                            crf + cf - cb - crb, crf + cf + cb + crb));
#endif
            }
            {  // y-face (vv.c:85-92)
                double av = vc;
                double uu = uk[L - 1 + RX], uru = uk[L + RX], ur = uc, ucc = uk[L - 1];
                double ulu = uk[L - 2 + RX], ul = uk[L - 2];
                double wu = wmu, wuf = wku, wf = wm1, wc = wmm;
                double cu = cc[L + RX], cru = cc[L + 1 + RX], cr = cc[L + 1];
                double cl = cc[L - 1], clu = cc[L - 1 + RX];
                double cuf = cq[L + RX], cf = c1p, cb = c1m, cub = cm[L + RX];
#ifdef MPD_FAST_DIV
                // (the divergence term's x strokes are shifted one cell left: uu, ucc shared)
                const double u1 = uu + ucc, w1 = wu + wc, w2 = wuf + wf;
                const double b1 = cru + cr, b2 = cl + clu, e1 = cuf + cf, e2 = cb + cub;
                double su = 0.25 * (u1 + (uru + ur));
                double sw = 0.25 * (w1 + w2);
                double div = 0.25 * av * ((vk[L + RX] - vk[L - RX]) + (w1 - w2) + (u1 - (ulu + ul)));
                put<NT>(a.Ay + zm1.off, colb,
                   antidiff(div, av, cu, c, su, b1 - b2, b1 + b2, sw, e1 - e2, e1 + e2));
#else
                double su = 0.25 * (uu + uru + ur + ucc);
                double sw = 0.25 * (wu + wuf + wf + wc);
                double div = 0.25 * av * (vk[L + RX] - vk[L - RX] + wu - wuf - wf + wc + uu - ulu -
                                          ul + ucc);
                put<NT>(a.Ay + zm1.off, colb,
                   antidiff(div, av, cu, c, su, cru + cr - cl - clu, cru + cr + cl + clu, sw,
                            cuf + cf - cb - cub, cuf + cf + cb + cub));
#endif
            }
            {  // z-face (ww.c:248-255)
                double av = wm1;
                double ur = uc, urf = un, uf = uq[L - 1], ucc = uk[L - 1];
                double vf = vq[L - RX], vuf = vn, vu = vc, vcc = vk[L - RX];
                double cf = c1p, crf = cq[L + 1], cr = cc[L + 1], cl = cc[L - 1];
                double clf = cq[L - 1], cuf = cq[L + RX], cu = cc[L + RX], cd = cc[L - RX];
                double cdf = cq[L - RX];
#ifdef MPD_FAST_DIV
                const double u1 = uf + ucc, u2 = urf + ur, v1 = vf + vcc, v2 = vuf + vu;
                const double b1 = crf + cr, b2 = cl + clf, e1 = cuf + cu, e2 = cd + cdf;
                double su = 0.25 * (u1 + u2);
                double sv = 0.25 * (v1 + v2);
                double div = 0.25 * av * ((wn - wmm) + (u1 - u2) + (v1 - v2));
                put<NT>(a.Az + zm1.off, colb,
                   antidiff(div, av, cf, c, su, b1 - b2, b1 + b2, sv, e1 - e2, e1 + e2));
#else
                double su = 0.25 * (ur + urf + uf + ucc);
                double sv = 0.25 * (vf + vuf + vu + vcc);
                double div = 0.25 * av * (wn - wmm + uf - urf - ur + ucc + vf - vuf - vu + vcc);
                put<NT>(a.Az + zm1.off, colb,
                   antidiff(div, av, cf, c, su, crf + cr - cl - clf, crf + cr + cl + clf, sv,
                            cuf + cu - cd - cdf, cuf + cu + cd + cdf));
#endif
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
        zp2.next();
        zp1.next();
        zp0.next();
        zm1.next();
    }
}

// F2 over one work item, one barrier per plane.  Iteration p: publish c^n(p), the pass
// input cp(p), the x/y antidiffusive faces of plane p and the limiter ratios of plane
// p - 1 (computed, on the tile and ring 1, in the previous iteration) -> barrier ->
// limiter ratios of plane p on the tile and ring 1 (held in registers until the next
// iteration publishes them) -> limited donor pass of plane k = p - 1 on the tile.
// LDS slots: c^n, Ax, Ay and the ratios two planes each (a fast wave publishes the next
// plane while slower waves still read this one; it cannot run further ahead than that
// before every wave has passed the next barrier), cp three (planes p and p - 1 are read
// while the next iteration writes p + 1).  The face values the donor pass needs a plane
// later are carried in registers.  SAVE also stores the ratios and the bound codes of
// every plane (the adjoint's forward record); it is a separate instantiation so the
// plain step does not carry its registers.
template <int TX, int TY, bool SAVE, bool NT>
__device__ __forceinline__ void f2_tile(const FusedArgs& a, int item)
{
    using G = Geo<TX, TY>;
    constexpr int RX = G::RX, R = G::R;
    __shared__ double s_cp[3][R], s_ax[2][R], s_ay[2][R], s_bu[2][R], s_bd[2][R];
    // bounds inputs: SAVE publishes c^n (its bound codes name the element of either field
    // that holds a bound); otherwise each cell publishes max(c^n, c_p) and min(c^n, c_p),
    // whose 7-point max / min are the same bounds in 12 instead of 26 operations
    __shared__ double s_cn[SAVE ? 2 : 1][SAVE ? R : 1], s_hi[SAVE ? 1 : 2][SAVE ? 1 : R],
        s_lo[SAVE ? 1 : 2][SAVE ? 1 : R];

    const int t = G::role(item);
    int bx, by, bz;
    tile_of(a, item, bx, by, bz);
    const int rx = 2 + t % TX, ry = 2 + t / TX;
    const int L = ry * RX + rx;
    unsigned colb = column<TX, TY>(a.nx, a.ny, bx, by, rx, ry);
    const bool store = bx * TX + rx - 2 < a.nx && by * TY + ry - 2 < a.ny;
    const bool hl = t < G::NH, h1 = t < G::H1;
    int hx = 0, hy = 0;
    if (hl) G::halo(t, hx, hy);
    const int HL = hy * RX + hx;
    unsigned hcolb = column<TX, TY>(a.nx, a.ny, bx, by, hx, hy);

    const int z0 = bz * a.zc;
    const int z1 = min(z0 + a.zc, a.nz);
    ZCursor zp2(z0 + 1, a.nz, a.np), zp1(z0, a.nz, a.np), zp0(z0 - 1, a.nz, a.np), zm1(z0 - 2, a.nz, a.np);

    // tile cell: c^n (p-1, p), cp (p-2 .. p), Az (p-2, p-1), Ax, Ay (p-1) with their
    // lower neighbours (L-1, L-RX), ratios (p-2, p-1); n*: loaded one iteration ahead
    double cnm = 0, cnc = 0, cpm2 = 0, cpm = 0, cpc = 0, azm2 = 0, azm = 0;
    double axm = 0, aym = 0, axlm = 0, aydm = 0;
    double bum2 = 0, bdm2 = 0, bum = 0, bdm = 0;
    double hbum = 0, hbdm = 0;  // halo cell: ratios of plane p - 1 (published in iteration p)
    double ncn = 0, ncp = 0, nax = 0, nay = 0, naz = 0;
    // halo cell
    double hcnm = 0, hcnc = 0, hcpm = 0, hcpc = 0, hazm = 0;
    double hncn = 0, hncp = 0, hnax = 0, hnay = 0, hnaz = 0;
    double nq = 0;  // emission of plane p, loaded in iteration p for plane k = p of p + 1
    cnm = ld(a.cn + zm1.off, colb);
    cnc = ld(a.cn + zp0.off, colb);
    cpm = ld(a.cp + zm1.off, colb);
    cpc = ld(a.cp + zp0.off, colb);
    azm = ld(a.Az + zm1.off, colb);
    ncn = ld(a.cn + zp1.off, colb);
    ncp = ld(a.cp + zp1.off, colb);
    nax = ld(a.Ax + zp0.off, colb);
    nay = ld(a.Ay + zp0.off, colb);
    naz = ld(a.Az + zp0.off, colb);
    if (hl) {
        hcnc = ld(a.cn + zp0.off, hcolb);
        hcpc = ld(a.cp + zp0.off, hcolb);
        hncn = ld(a.cn + zp1.off, hcolb);
        hncp = ld(a.cp + zp1.off, hcolb);
        hnax = ld(a.Ax + zp0.off, hcolb);
        hnay = ld(a.Ay + zp0.off, hcolb);
        if (h1) {
            hcnm = ld(a.cn + zm1.off, hcolb);
            hcpm = ld(a.cp + zm1.off, hcolb);
            hazm = ld(a.Az + zm1.off, hcolb);
            hnaz = ld(a.Az + zp0.off, hcolb);
        }
    }
    // the plane loop, unrolled by hand (four planes per trip, then two, then an odd last
    // plane): across the copies the compiler renames the registers the planes below and
    // above rotate through instead of copying them (the barriers inside keep it from
    // unrolling a loop of unknown trip count itself).  The SAVE variant, held to three
    // waves per SIMD, would spill with four copies and keeps two.
    auto plane = [&](const int pl) __attribute__((always_inline)) {
        const int p = __builtin_amdgcn_readfirstlane(pl);  // uniform, as in f1_tile
        keep32(colb);
        keep32(hcolb);
        const int s2 = (p + 2) % 2, s3 = (p + 3) % 3;
        __builtin_amdgcn_s_setprio(2);  // until the barrier, as in f1_tile
        const double cnn = ncn, cpn = ncp, axn = nax, ayn = nay, azc = naz;
        const double hcnn = hncn, hcpn = hncp, haxn = hnax, hayn = hnay, hazc = hnaz;
        const double qk = nq;
        if (a.q && p >= z0 && p < z1) nq = ld(a.q + zp0.off, colb);
        {
            // the next iteration's values: fields at p + 2, faces at p + 1.  Loaded by every
            // lane in every iteration (the last one reloads planes already in the cache, a
            // lane without a halo cell the wrapped column of region cell (0, 0)): loads
            // under a condition kept the previous values alive beside the new ones and
            // cost ~15 register copies per plane.  Fresh copies of the column offsets keep
            // the loads in the scalar-base form.
            const bool more = p < z1;
            const size_t o2 = more ? zp2.off : zp1.off, o1 = more ? zp1.off : zp0.off;
            unsigned c2 = colb, h2 = hcolb;
            keep32(c2);
            keep32(h2);
            ncn = ld(a.cn + o2, c2);
            ncp = ld(a.cp + o2, c2);
            nax = ld(a.Ax + o1, c2);
            nay = ld(a.Ay + o1, c2);
            naz = ld(a.Az + o1, c2);
            hncn = ld(a.cn + o2, h2);
            hncp = ld(a.cp + o2, h2);
            hnax = ld(a.Ax + o1, h2);
            hnay = ld(a.Ay + o1, h2);
            hnaz = ld(a.Az + o1, h2);
        }
        const int k2 = (p + 1) % 2;  // slot of plane p - 1
        if constexpr (SAVE) {
            s_cn[s2][L] = cnc;
        } else {
            s_hi[s2][L] = mx2(cnc, cpc);
            s_lo[s2][L] = mn2(cnc, cpc);
        }
        s_cp[s3][L] = cpc;
        s_ax[s2][L] = axn;
        s_ay[s2][L] = ayn;
        s_bu[k2][L] = bum;
        s_bd[k2][L] = bdm;
        if (hl) {
            if constexpr (SAVE) {
                s_cn[s2][HL] = hcnc;
            } else {
                s_hi[s2][HL] = mx2(hcnc, hcpc);
                s_lo[s2][HL] = mn2(hcnc, hcpc);
            }
            s_cp[s3][HL] = hcpc;
            s_ax[s2][HL] = haxn;
            s_ay[s2][HL] = hayn;
        }
        if (h1) {
            s_bu[k2][HL] = hbum;
            s_bd[k2][HL] = hbdm;
        }
        lds_barrier();
        __builtin_amdgcn_s_setprio(0);
        // limiter ratios of plane p (see cpu_ref.c)
        const double* cpp = s_cp[s3];
        auto ratios = [&](int M, double cn0, double cnlo, double cnhi, double c, double cplo,
                          double cphi, double axr, double ayr, double azlo, double azhi,
                          double axl, double ayl, double& bu, double& bd, int* code) {
            const double nbp[7] = {c, cpp[M - 1], cpp[M + 1], cpp[M - RX], cpp[M + RX], cplo,
                                   cphi};
            double hi, lo;
            if constexpr (!SAVE) {
                const double* hp = s_hi[s2];
                const double* lp = s_lo[s2];
                hi = mx2(mx2(mx2(mx2(cn0, c), hp[M - 1]), mx2(hp[M + 1], hp[M - RX])),
                         mx2(mx2(hp[M + RX], mx2(cnlo, cplo)), mx2(cnhi, cphi)));
                lo = mn2(mn2(mn2(mn2(cn0, c), lp[M - 1]), mn2(lp[M + 1], lp[M - RX])),
                         mn2(mn2(lp[M + RX], mn2(cnlo, cplo)), mn2(cnhi, cphi)));
            } else {
                hi = lo = cn0;
                const double* cnp = s_cn[s2];
                const double nbn[7] = {cn0,         cnp[M - 1], cnp[M + 1], cnp[M - RX],
                                       cnp[M + RX], cnlo,       cnhi};
#pragma unroll
                for (int m = 0; m < 7; m++) {
                    hi = mx2(hi, mx2(nbn[m], nbp[m]));
                    lo = mn2(lo, mn2(nbn[m], nbp[m]));
                }
                if (code) {
                    // which element holds each bound: element 2m + s is nbn[m] (s = 0) or
                    // nbp[m] (s = 1), and of the elements equal to the bound the last one
                    // counts (the convention of a running maximum that replaces on ties)
                    int hc = 0, lc = 0;
#pragma unroll
                    for (int m = 0; m < 7; m++) {
                        hc = nbn[m] == hi ? 2 * m : hc;
                        lc = nbn[m] == lo ? 2 * m : lc;
                        hc = nbp[m] == hi ? 2 * m + 1 : hc;
                        lc = nbp[m] == lo ? 2 * m + 1 : lc;
                    }
                    *code = hc | (lc << 8);
                }
            }
            const double fxl = 0.5 * F2(nbp[1], c, axl), fxr = 0.5 * F2(c, nbp[2], axr);
            const double fyl = 0.5 * F2(nbp[3], c, ayl), fyr = 0.5 * F2(c, nbp[4], ayr);
            const double fzl = 0.5 * F2(cplo, c, azlo), fzr = 0.5 * F2(c, cphi, azhi);
            beta(hi, lo, c, fxl, fxr, fyl, fyr, fzl, fzr, bu, bd);
            if constexpr (!SAVE) {  // clamped once here (SAVE records the ratios themselves)
                bu = mn2(bu, 1.0);
                bd = mn2(bd, 1.0);
            }
        };
        const double axl = s_ax[s2][L - 1], ayd = s_ay[s2][L - RX];
        double buc, bdc;
        int code = 0;
        ratios(L, cnc, cnm, cnn, cpc, cpm, cpn, axn, ayn, azm, azc, axl, ayd, buc, bdc,
               SAVE ? &code : nullptr);
        if (SAVE && store && p >= z0 && p < z1) {
            put<NT>(a.bu + zp0.off, colb, buc);
            put<NT>(a.bd + zp0.off, colb, bdc);
            int* cp_ = (int*)((char*)(a.code + zp0.off) + colb / 2);
            if constexpr (NT)
                __builtin_nontemporal_store(code, cp_);
            else
                *cp_ = code;
        }
        if (h1)  // (the plane p - 1 values are dead: published above)
            ratios(HL, hcnc, hcnm, hcnn, hcpc, hcpm, hcpn, haxn, hayn, hazm, hazc,
                   s_ax[s2][HL - 1], s_ay[s2][HL - RX], hbum, hbdm, nullptr);
        const int k = p - 1;
        if (k >= z0 && k < z1) {
            const int k3 = (p + 2) % 3;  // slot of plane k (c_p; the ratios: k2)
            const double* bu = s_bu[k2];
            const double* bd = s_bd[k2];
            const double* cq = s_cp[k3];
            const double b0u = bum, b0d = bdm;
            auto lim = [](double u, double bul, double bdl, double bur, double bdr) {
                return SAVE ? limit_face(u, bul, bdl, bur, bdr)
                            : limit_face_clamped(u, bul, bdl, bur, bdr);
            };
            const double uxr = lim(axm, b0u, b0d, bu[L + 1], bd[L + 1]);
            const double uxl = lim(axlm, bu[L - 1], bd[L - 1], b0u, b0d);
            const double vyr = lim(aym, b0u, b0d, bu[L + RX], bd[L + RX]);
            const double vyl = lim(aydm, bu[L - RX], bd[L - RX], b0u, b0d);
            const double wzr = lim(azm, b0u, b0d, buc, bdc);
            const double wzl = lim(azm2, bum2, bdm2, b0u, b0d);
            double r = donor(cpm, cq[L - 1], cq[L + 1], cq[L - RX], cq[L + RX], cpm2, cpc, uxl,
                             uxr, vyl, vyr, wzl, wzr);
            if (store) {
                if (a.q) r = r + qk;
                put<NT>(a.cout + zm1.off, colb, r);
                if (a.Lx) {
                    put<NT>(a.Lx + zm1.off, colb, uxr);
                    put<NT>(a.Ly + zm1.off, colb, vyr);
                    put<NT>(a.Lz + zm1.off, colb, wzr);
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
        axm = axn;
        aym = ayn;
        axlm = axl;
        aydm = ayd;
        bum2 = bum;
        bdm2 = bdm;
        bum = buc;
        bdm = bdc;
        hcnm = hcnc;
        hcnc = hcnn;
        hcpm = hcpc;
        hcpc = hcpn;
        hazm = hazc;
        zp2.next();
        zp1.next();
        zp0.next();
        zm1.next();
    };
    int pl = z0 - 1;
    if constexpr (!SAVE) {
        for (; pl + 3 <= z1; pl += 4) {
            plane(pl);
            plane(pl + 1);
            plane(pl + 2);
            plane(pl + 3);
        }
    }
    for (; pl + 1 <= z1; pl += 2) {
        plane(pl);
        plane(pl + 1);
    }
    if (pl <= z1) plane(pl);
}

// F1 is held to 128 registers (4 waves per SIMD; a 16-byte spill in the donor variant),
// which measures faster than the 3 waves its natural 144 registers allow.  F2 would spill
// at that bound and keeps its natural allocation (3 waves per SIMD); its SAVE variant is
// held to the same 3 waves (168 registers), below which its natural allocation would
// fall, for 4-wave tiles (8-wave tiles are held to one workgroup per CU by LDS anyway).
template <int TX, int TY, bool DONOR, bool NT>
__global__ void __launch_bounds__((Geo<TX, TY>::NT)) __attribute__((amdgpu_waves_per_eu(4)))
k_f1(FusedArgs a, Work w)
{
    for (int it = blockIdx.x; it < w.nitems; it += gridDim.x) {
        int item = item_of(w, it);
        FusedArgs x = a;
        if (w.ninst) {
            ConstInt* item0 = (ConstInt*)w.item0;
            int i = find_inst(item0, w.ninst, item);
            __builtin_memcpy(&x, (const FusedArgs*)((ConstArgs*)w.table + i), sizeof x);
            item -= item0[i];
        }
        f1_tile<TX, TY, DONOR, NT>(x, item);
        lds_barrier();  // the next item reuses the LDS planes
    }
}

template <int TX, int TY, bool SAVE, bool NT>
__global__ void __launch_bounds__((Geo<TX, TY>::NT)) __attribute__((amdgpu_waves_per_eu(Geo<TX, TY>::NT <= 256 ? 3 : 1)))
k_f2(FusedArgs a, Work w)
{
    for (int it = blockIdx.x; it < w.nitems; it += gridDim.x) {
        int item = item_of(w, it);
        FusedArgs x = a;
        if (w.ninst) {
            ConstInt* item0 = (ConstInt*)w.item0;
            int i = find_inst(item0, w.ninst, item);
            __builtin_memcpy(&x, (const FusedArgs*)((ConstArgs*)w.table + i), sizeof x);
            item -= item0[i];
        }
        f2_tile<TX, TY, SAVE, NT>(x, item);
        lds_barrier();
    }
}

// The donor pass of the unlimited corrective pass, one thread per cell on a 3-D grid of
// 64 x 4 (x, y) blocks and one z plane per blockIdx.z (plane k0 + blockIdx.z): indices
// from block coordinates, no 64-bit division.
constexpr int DBX = 64, DBY = 4, DMAXZ = 65535;
__global__ void __launch_bounds__(DBX * DBY)
k_donor_faces(int nx, int ny, int nz, int k0, const double* __restrict__ U,
              const double* __restrict__ V, const double* __restrict__ W,
              const double* __restrict__ cin, const double* __restrict__ q,
              double* __restrict__ cout)
{
    const size_t np = (size_t)nx * ny;
    const int i = blockIdx.x * DBX + threadIdx.x, j = blockIdx.y * DBY + threadIdx.y;
    const int k = k0 + (int)blockIdx.z;
    if (i >= nx || j >= ny) return;
    const size_t t = (size_t)k * np + (size_t)j * nx + i;
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

// ------------------------------------------------------ one kernel per pass ---
//
// F12: a whole corrective pass in one kernel -- [donor field c1 = donor(c^n)],
// antidiffusive velocities A of c1, limiter ratios, limited donor pass -- so that c1 and
// A never leave the chip.  The price is recomputing them on a wider halo: the ratios on
// ring 1 of the tile need A on the faces of ring-1 cells, which needs c1 on ring 2,
// which needs c^n (and the velocities) on ring 3.
//
// Geometry: a TX x TY tile with a three-cell ring (the region).  The workgroup's waves
// split the columns by role: TX*TY/64 tile waves own one tile column per thread; four
// ring waves own the ring columns, each lane at most one cell of rings 1-2 (the primary:
// c1, and on ring 1 and the ring-2 cells left of / below ring 1 the faces the tile needs)
// and at most one cell of ring 3 (the secondary: loads and publishes only).  Ring 1 and
// ring 2 are dealt evenly over the four ring waves, so every SIMD (three waves: two tile
// waves and one ring wave) carries the same work per plane, and no thread holds the
// rolling state of two computed columns -- which keeps a wave within the 168 registers
// that three waves per SIMD allow (the LDS planes below hold a CU to one workgroup).
//
// Plane pipeline, iteration p (two barriers):
//   phase 0  publish c^n(p), U(p), V(p); prefetch the next planes
//   phase 1  c1(p) on the tile and rings 1-2; publish it with W(p-1)
//   phase 2  A of plane k = p-1 (x/y faces of k, z-face k|k+1) on the tile and ring 1;
//            ratios of plane m = p-2 on the tile and ring 1; the limited donor pass of
//            plane o = p-3 on the tile.
// LDS slots per quantity follow from how far a fast wave can run ahead (one phase):
// c^n 4 (read for the ratios two iterations after it is published), U and V 3, c1 4,
// W 1, A (x, y) 2, ratios 2.
template <int TX, int TY>
struct Geo3 {
    static constexpr int RX = TX + 6, RY = TY + 6, R = RX * RY;
    static constexpr int NT = TX * TY;
    static constexpr int H1 = 2 * (TX + 2) + 2 * TY;      // ring 1
    static constexpr int H2A = TY + 2, H2B = TX + 2;      // ring-2 cells with one face
    static constexpr int H2 = 2 * (TX + 4) + 2 * (TY + 2);  // ring 2
    static constexpr int H3 = 2 * (TX + 6) + 2 * (TY + 4);  // ring 3
    static constexpr int NH = H1 + H2 + H3;
    // ring waves and the cells per ring wave of ring 1, ring 2 (primary) and ring 3
    static constexpr int NRW = 4;
    static constexpr int Q1 = (H1 + NRW - 1) / NRW, Q2 = (H2 + NRW - 1) / NRW;
    static constexpr int Q3 = (H3 + NRW - 1) / NRW;
    static constexpr int NTH = NT + 64 * NRW;  // threads of the workgroup
    static_assert(Q1 + Q2 <= 64 && Q3 <= 64, "a ring lane owns at most one cell of each kind");
    static_assert(NT % 64 == 0, "whole tile waves");
    // Ring cell (halo numbering below) of ring lane l as primary / secondary, or -1.
    __device__ static int primary(int l)
    {
        const int r = l >> 6, j = l & 63;
        if (j < Q1) return r * Q1 + j < H1 ? r * Q1 + j : -1;
        if (j < Q1 + Q2) return r * Q2 + j - Q1 < H2 ? H1 + r * Q2 + j - Q1 : -1;
        return -1;
    }
    __device__ static int secondary(int l)
    {
        const int r = l >> 6, j = l & 63;
        return j < Q3 && r * Q3 + j < H3 ? H1 + H2 + r * Q3 + j : -1;
    }
    // Region cell (rx, ry) of ring cell h and its kind: 1 ring 1, 2 ring 2 left of ring 1
    // (x-face), 3 ring 2 below ring 1 (y-face), 4 the other ring-2 cells, 5 ring 3.
    __device__ static int halo(int h, int& rx, int& ry)
    {
        if (h < H1) {
            if (h < 2 * (TX + 2)) {
                ry = h < TX + 2 ? 2 : TY + 3;
                rx = 2 + h % (TX + 2);
            } else {
                const int r = h - 2 * (TX + 2);
                ry = 3 + r / 2;
                rx = (r & 1) ? TX + 3 : 2;
            }
            return 1;
        }
        h -= H1;
        if (h < H2A) {
            rx = 1;
            ry = 2 + h;
            return 2;
        }
        h -= H2A;
        if (h < H2B) {
            ry = 1;
            rx = 2 + h;
            return 3;
        }
        h -= H2B;
        if (h < H2 - H2A - H2B) {
            if (h < 4) {
                rx = (h & 1) ? TX + 4 : 1;
                ry = (h & 2) ? TY + 4 : 1;
            } else if (h < 4 + TY + 2) {
                rx = TX + 4;
                ry = 2 + (h - 4);
            } else {
                ry = TY + 4;
                rx = 2 + (h - 4 - (TY + 2));
            }
            return 4;
        }
        h -= H2 - H2A - H2B;
        if (h < 2 * RX) {
            ry = h < RX ? 0 : RY - 1;
            rx = h % RX;
        } else {
            const int r = h - 2 * RX;
            ry = 1 + r / 2;
            rx = (r & 1) ? RX - 1 : 0;
        }
        return 5;
    }
};

// The LDS planes of F12 (one region per slot).
template <int R>
struct F12Lds {
    double cn[4][R], u[3][R], v[3][R], w[R], c1[4][R], ax[2][R], ay[2][R], bu[2][R], bd[2][R];
};

template <int TX, int TY>
__device__ __forceinline__ unsigned column3(int nx, int ny, int bx, int by, int cx, int cy)
{
    const int gx = wrapi(bx * TX + cx - 3, nx), gy = wrapi(by * TY + cy - 3, ny);
    return (unsigned)(gy * nx + gx) * 8u;
}

// Rolling per-column state of F12 (a tile column or a ring column).
struct Col3 {
    unsigned colb;
    double cn_m3 = 0, cn_m2 = 0, cn_m1 = 0, cn_0 = 0, cn_p1 = 0, ncn = 0;  // c^n(p-3 .. p+2)
    double w_m2 = 0, w_m1 = 0, w_0 = 0, nw = 0;                           // W(p-2 .. p+1)
    double u_m1 = 0, u_0 = 0, nu = 0, v_m1 = 0, v_0 = 0, nv = 0;          // U, V(p-1 .. p+1)
    double fld = 0, nf = 0;                                               // pass input (p, p+1)
    double wr_m = 0, wu_m = 0;        // W of the x / y neighbour at plane p-2
    double c1_m4 = 0, c1_m3 = 0, c1_m2 = 0, c1_m1 = 0, c1_0 = 0;         // c1(p-4 .. p)
    double ax_m3 = 0, ay_m3 = 0, ax_m2 = 0, ay_m2 = 0;                   // own A_x, A_y
    double az_m4 = 0, az_m3 = 0, az_m2 = 0;                              // own A_z
    double axl_m3 = 0, ayd_m3 = 0;    // A_x of the x-, A_y of the y- neighbour, plane p-3
    double bu_m4 = 0, bd_m4 = 0, bu_m3 = 0, bd_m3 = 0;                   // own ratios
};

// One role's share of a work item.  TILE: lane = tile cell (all stages, the stores).
// !TILE: lane = ring lane, with its primary column (rings 1-2: c1; the faces and ratios
// the tile reads) and its secondary column (ring 3: c^n, U, V for the ring-2 donor field).
// SAVE (first pass of an iord-2 step only): the tile also stores the adjoint's forward
// record of every tile cell -- c1, the unlimited faces A, the unclamped limiter ratios and
// the bound codes (see FusedArgs) -- the values the two-kernel storing step records, bit
// for bit.
template <int TX, int TY, bool DONOR, bool NT, bool SAVE, bool EMIT, bool TILE>
__device__ __forceinline__ void f12_part(const FusedArgs& a, int item, int lane,
                                         F12Lds<Geo3<TX, TY>::R>& sm)
{
    using G = Geo3<TX, TY>;
    constexpr int RX = G::RX;
    int bx, by, bz;
    tile_of(a, item, bx, by, bz);
    // the primary column (rx, ry) and its kind (0: none; the tile's cells count as kind 1
    // with every stage); the secondary (sx, sy) of a ring lane, if any
    int rx = 3, ry = 3, sx = 3, sy = 3, kind = 1;
    bool sec = false;
    if constexpr (TILE) {
        rx = 3 + lane % TX;
        ry = 3 + lane / TX;
    } else {
        const int hp = G::primary(lane), hs = G::secondary(lane);
        kind = hp >= 0 ? G::halo(hp, rx, ry) : 0;
        if (hs >= 0) {
            G::halo(hs, sx, sy);
            sec = true;
        }
    }
    const int L = ry * RX + rx, LS = sy * RX + sx;
    const bool act = TILE || kind != 0;                   // loads, c1, publishes
    const bool fxk = TILE || kind == 1 || kind == 2;      // the face right of the cell
    const bool fyk = TILE || kind == 1 || kind == 3;      // the face above it
    const bool r1 = TILE || kind == 1;                    // z-face and ratios
    const bool store = TILE && bx * TX + rx - 3 < a.nx && by * TY + ry - 3 < a.ny;
    Col3 T, S;
    T.colb = column3<TX, TY>(a.nx, a.ny, bx, by, rx, ry);
    S.colb = column3<TX, TY>(a.nx, a.ny, bx, by, sx, sy);

    const int z0 = bz * a.zc;
    const int z1 = min(z0 + a.zc, a.nz);
    const int pstart = z0 - 2, pend = z1 + 2;
    // planes p+2, p+1 (loads), p-2 (emission of the next output plane), p-3 (stores)
    ZCursor zp2(pstart + 2, a.nz, a.np), zp1(pstart + 1, a.nz, a.np),
        zm2(pstart - 2, a.nz, a.np), zm3(pstart - 3, a.nz, a.np);
    ZCursor zp0(pstart, a.nz, a.np), zm1(pstart - 1, a.nz, a.np);  // records (SAVE)
    // Every lane loads, idle ring lanes at a valid column of the region: a load under a
    // divergent condition makes the compiler copy its value out at the join, which waits
    // for the load -- once per plane, a full memory round trip.
    {
        ZCursor zb(pstart - 1, a.nz, a.np), zc(pstart, a.nz, a.np);
        T.cn_m1 = ld(a.cn + zb.off, T.colb);
        T.cn_0 = ld(a.cn + zc.off, T.colb);
        if (!DONOR) T.nf = ld(a.cin + zc.off, T.colb);
        T.ncn = ld(a.cn + zp1.off, T.colb);
        T.w_m1 = ld(a.W + zb.off, T.colb);
        T.nw = ld(a.W + zc.off, T.colb);
        T.nu = ld(a.U + zc.off, T.colb);
        T.nv = ld(a.V + zc.off, T.colb);
        if constexpr (!TILE) {
            S.cn_0 = ld(a.cn + zc.off, S.colb);
            S.ncn = ld(a.cn + zp1.off, S.colb);
            S.nu = ld(a.U + zc.off, S.colb);
            S.nv = ld(a.V + zc.off, S.colb);
        }
    }
    double nq = 0;  // emission of plane p-2, used for the output plane of the next iteration
    // (an instance of a batch without emission reads a plane of its field instead: an
    // unconditional load, whatever the item)
    const double* qb = a.q ? a.q : a.cn;

    // the plane loop, unrolled by hand (four planes per trip, then one at a time): across
    // the copies the per-column registers are renamed instead of copied (the barriers
    // inside keep the compiler from unrolling a loop of unknown trip count itself)
    // steady: every stage is active (planes z0 + 3 .. z1 + 1), so the body has no stage
    // guards and the compiler may move each stage's LDS reads ahead of the previous
    // stage's arithmetic
    auto plane = [&](const int pl, auto steady) __attribute__((always_inline)) {
        constexpr bool ST = decltype(steady)::value;
        const int p = __builtin_amdgcn_readfirstlane(pl);
        keep32(T.colb);
        if constexpr (!TILE) keep32(S.colb);
        const int p8 = p + 8;  // p >= -2: non-negative slot arithmetic
        const int s4 = p8 & 3, s3 = p8 % 3;
        // ---- phase 0: this plane's inputs, the next planes' loads
        __builtin_amdgcn_s_setprio(2);
        const double qo = nq;
        if constexpr (TILE && EMIT) nq = ld(qb + zm2.off, T.colb);
        T.cn_p1 = T.ncn;
        T.u_0 = T.nu;
        T.v_0 = T.nv;
        T.w_0 = T.nw;
        T.fld = T.nf;
        {
            unsigned c2 = T.colb;
            keep32(c2);
            T.ncn = ld(a.cn + zp2.off, c2);
            T.nu = ld(a.U + zp1.off, c2);
            T.nv = ld(a.V + zp1.off, c2);
            T.nw = ld(a.W + zp1.off, c2);
            if (!DONOR) T.nf = ld(a.cin + zp1.off, c2);
        }
        if (act) {
            sm.cn[s4][L] = T.cn_0;
            sm.u[s3][L] = T.u_0;
            sm.v[s3][L] = T.v_0;
        }
        if constexpr (!TILE) {
            S.u_0 = S.nu;
            S.v_0 = S.nv;
            unsigned c2 = S.colb;
            keep32(c2);
            const double cn = S.cn_0;
            S.cn_0 = S.ncn;  // (c^n(p+1): the secondary keeps no plane below p)
            S.ncn = ld(a.cn + zp2.off, c2);
            S.nu = ld(a.U + zp1.off, c2);
            S.nv = ld(a.V + zp1.off, c2);
            if (sec) {
                sm.cn[s4][LS] = cn;
                sm.u[s3][LS] = S.u_0;
                sm.v[s3][LS] = S.v_0;
            }
        }
        lds_barrier();
        // the ring waves' work gates the next plane of every wave: they take precedence
        // over the (older) tile waves, which fill the cycles the ring waves leave
        __builtin_amdgcn_s_setprio(TILE ? 0 : 1);
        // ---- phase 1: c1 of plane p
        const bool do_c1 = ST || p <= z1 + 1;
        if (act) {
            if (do_c1) {
                const double* cn = sm.cn[s4];
                const double* up = sm.u[s3];
                const double* vp = sm.v[s3];
                T.c1_0 = DONOR ? donor(T.cn_0, cn[L - 1], cn[L + 1], cn[L - RX], cn[L + RX],
                                       T.cn_m1, T.cn_p1, up[L - 1], T.u_0, vp[L - RX], T.v_0,
                                       T.w_m1, T.w_0)
                               : T.fld;
                sm.c1[s4][L] = T.c1_0;
                if constexpr (SAVE && TILE)
                    if (store && p >= z0 && p < z1) put<NT>(a.c1out + zp0.off, T.colb, T.c1_0);
            }
            sm.w[L] = T.w_m1;
        }
        lds_barrier();
        // ---- phase 2: faces, ratios and (tile) the limited donor pass
        const int k = p - 1, m = p - 2, o = p - 3;
        const bool doA = ST || (k >= z0 - 2 && k <= z1);
        const bool doB = ST || (m >= z0 - 1 && m <= z1);
        const bool doO = ST || (o >= z0 && o < z1);
        const int km4 = (p8 - 1) & 3, km3 = (p8 - 1) % 3, kq4 = s4, kq3 = s3;  // planes k, k+1
        const int kb4 = (p8 - 2) & 3;                                          // plane k-1
        const int ka = (p8 - 1) & 1;
        // (ring cells that compute no face are not read at their neighbours' positions,
        // some of which lie outside the region)
        const double wkr = sm.w[fxk ? L + 1 : L], wku = sm.w[fyk ? L + RX : L];
        const double* cm = sm.c1[kb4];
        const double* cc = sm.c1[km4];
        const double* cq = sm.c1[kq4];
        const double* uk = sm.u[km3];
        const double* uq = sm.u[kq3];
        const double* vk = sm.v[km3];
        const double* vq = sm.v[kq3];
        // antidiffusive velocities of the faces above/right of cell M of plane k
        // (the formulas and operand order of f1_tile)
        auto faces = [&](int M, const Col3& C, double wr, double wu, bool fx, bool fy,
                         bool fz, double& ox, double& oy, double& oz)
                         __attribute__((always_inline)) {
            const double c = C.c1_m1;
            if (fx) {  // x-face (uu.c:81-87)
                double av = C.u_m1;
                double vr = vk[M + 1 - RX], vru = vk[M + 1], vu = C.v_m1, vcc = vk[M - RX];
                double wr_ = C.wr_m, wrf = wr, wf = C.w_m1, wc = C.w_m2;
                double cr = cc[M + 1], cru = cc[M + 1 + RX], cu = cc[M + RX];
                double cd = cc[M - RX], crd = cc[M + 1 - RX];
                double crf = cq[M + 1], cf = C.c1_0, cb = C.c1_m2, crb = cm[M + 1];
#ifdef MPD_FAST_DIV
                const double v1 = vr + vcc, v2 = vru + vu, w1 = wr_ + wc, w2 = wrf + wf;
                const double b1 = cru + cu, b2 = cd + crd, e1 = crf + cf, e2 = cb + crb;
                double sv = 0.25 * (v1 + v2);
                double sw = 0.25 * (w1 + w2);
                double div = 0.25 * av * ((uk[M + 1] - uk[M - 1]) + (v1 - v2) + (w1 - w2));
                ox = antidiff(div, av, cr, c, sv, b1 - b2, b1 + b2, sw, e1 - e2, e1 + e2);
#else
                double sv = 0.25 * (vr + vru + vu + vcc);
                double sw = 0.25 * (wr_ + wrf + wf + wc);
                double div = 0.25 * av * (uk[M + 1] - uk[M - 1] + vr - vru - vu + vcc + wr_ -
                                          wrf - wf + wc);
                ox = antidiff(div, av, cr, c, sv, cru + cu - cd - crd, cru + cu + cd + crd,
                              sw, crf + cf - cb - crb, crf + cf + cb + crb);
#endif
            }
            if (fy) {  // y-face (vv.c:85-92)
                double av = C.v_m1;
                double uu = uk[M - 1 + RX], uru = uk[M + RX], ur = C.u_m1, ucc = uk[M - 1];
                double ulu = uk[M - 2 + RX], ul = uk[M - 2];
                double wu_ = C.wu_m, wuf = wu, wf = C.w_m1, wc = C.w_m2;
                double cu = cc[M + RX], cru = cc[M + 1 + RX], cr = cc[M + 1];
                double cl = cc[M - 1], clu = cc[M - 1 + RX];
                double cuf = cq[M + RX], cf = C.c1_0, cb = C.c1_m2, cub = cm[M + RX];
#ifdef MPD_FAST_DIV
                const double u1 = uu + ucc, w1 = wu_ + wc, w2 = wuf + wf;
                const double b1 = cru + cr, b2 = cl + clu, e1 = cuf + cf, e2 = cb + cub;
                double su = 0.25 * (u1 + (uru + ur));
                double sw = 0.25 * (w1 + w2);
                double div =
                    0.25 * av * ((vk[M + RX] - vk[M - RX]) + (w1 - w2) + (u1 - (ulu + ul)));
                oy = antidiff(div, av, cu, c, su, b1 - b2, b1 + b2, sw, e1 - e2, e1 + e2);
#else
                double su = 0.25 * (uu + uru + ur + ucc);
                double sw = 0.25 * (wu_ + wuf + wf + wc);
                double div = 0.25 * av * (vk[M + RX] - vk[M - RX] + wu_ - wuf - wf + wc + uu -
                                          ulu - ul + ucc);
                oy = antidiff(div, av, cu, c, su, cru + cr - cl - clu, cru + cr + cl + clu,
                              sw, cuf + cf - cb - cub, cuf + cf + cb + cub);
#endif
            }
            if (fz) {  // z-face (ww.c:248-255)
                double av = C.w_m1;
                double ur = C.u_m1, urf = C.u_0, uf = uq[M - 1], ucc = uk[M - 1];
                double vf = vq[M - RX], vuf = C.v_0, vu = C.v_m1, vcc = vk[M - RX];
                double cf = C.c1_0, crf = cq[M + 1], cr = cc[M + 1], cl = cc[M - 1];
                double clf = cq[M - 1], cuf = cq[M + RX], cu = cc[M + RX], cd = cc[M - RX];
                double cdf = cq[M - RX];
#ifdef MPD_FAST_DIV
                const double u1 = uf + ucc, u2 = urf + ur, v1 = vf + vcc, v2 = vuf + vu;
                const double b1 = crf + cr, b2 = cl + clf, e1 = cuf + cu, e2 = cd + cdf;
                double su = 0.25 * (u1 + u2);
                double sv = 0.25 * (v1 + v2);
                double div = 0.25 * av * ((C.w_0 - C.w_m2) + (u1 - u2) + (v1 - v2));
                oz = antidiff(div, av, cf, c, su, b1 - b2, b1 + b2, sv, e1 - e2, e1 + e2);
#else
                double su = 0.25 * (ur + urf + uf + ucc);
                double sv = 0.25 * (vf + vuf + vu + vcc);
                double div = 0.25 * av * (C.w_0 - C.w_m2 + uf - urf - ur + ucc + vf - vuf -
                                          vu + vcc);
                oz = antidiff(div, av, cf, c, su, crf + cr - cl - clf, crf + cr + cl + clf,
                              sv, cuf + cu - cd - cdf, cuf + cu + cd + cdf);
#endif
            }
        };
        // limiter ratios of plane m (see cpu_ref.c), clamped to at most 1
        const int mb4 = (p8 - 2) & 3, ma = (p8 - 2) & 1;
        const double* cnm = sm.cn[mb4];
        const double* c1m = sm.c1[mb4];
        const double* axm = sm.ax[ma];
        const double* aym = sm.ay[ma];
        auto ratios = [&](int M, const Col3& C, double& bu, double& bd, double& axl,
                          double& ayd) __attribute__((always_inline)) {
            const double c = C.c1_m2;
            const double xl = c1m[M - 1], xr = c1m[M + 1], yd = c1m[M - RX], yu = c1m[M + RX];
            const double hi =
                mx2(mx2(mx2(mx2(C.cn_m2, c), mx2(cnm[M - 1], xl)), mx2(mx2(cnm[M + 1], xr),
                                                                      mx2(cnm[M - RX], yd))),
                    mx2(mx2(mx2(cnm[M + RX], yu), mx2(C.cn_m3, C.c1_m3)),
                        mx2(C.cn_m1, C.c1_m1)));
            const double lo =
                mn2(mn2(mn2(mn2(C.cn_m2, c), mn2(cnm[M - 1], xl)), mn2(mn2(cnm[M + 1], xr),
                                                                      mn2(cnm[M - RX], yd))),
                    mn2(mn2(mn2(cnm[M + RX], yu), mn2(C.cn_m3, C.c1_m3)),
                        mn2(C.cn_m1, C.c1_m1)));
            axl = axm[M - 1];
            ayd = aym[M - RX];
            const double fxl = 0.5 * F2(xl, c, axl), fxr = 0.5 * F2(c, xr, C.ax_m2);
            const double fyl = 0.5 * F2(yd, c, ayd), fyr = 0.5 * F2(c, yu, C.ay_m2);
            const double fzl = 0.5 * F2(C.c1_m3, c, C.az_m3), fzr = 0.5 * F2(c, C.c1_m1, C.az_m2);
            beta(hi, lo, c, fxl, fxr, fyl, fyr, fzl, fzr, bu, bd);
            if constexpr (SAVE && TILE) {
                if (store && m >= z0 && m < z1) {
                    // which element holds each bound: element 2q + s is the c^n (s = 0) or
                    // pass-input (s = 1) value of neighbour q in {self, x-, x+, y-, y+, z-,
                    // z+}; of the elements equal to the bound the last one counts
                    const double nbn[7] = {C.cn_m2,     cnm[M - 1], cnm[M + 1], cnm[M - RX],
                                           cnm[M + RX], C.cn_m3,    C.cn_m1};
                    const double nbp[7] = {c, xl, xr, yd, yu, C.c1_m3, C.c1_m1};
                    int hc = 0, lc = 0;
#pragma unroll
                    for (int q = 0; q < 7; q++) {
                        hc = nbn[q] == hi ? 2 * q : hc;
                        lc = nbn[q] == lo ? 2 * q : lc;
                        hc = nbp[q] == hi ? 2 * q + 1 : hc;
                        lc = nbp[q] == lo ? 2 * q + 1 : lc;
                    }
                    put<NT>(a.bu + zm2.off, T.colb, bu);
                    put<NT>(a.bd + zm2.off, T.colb, bd);
                    int* cp_ = (int*)((char*)(a.code + zm2.off) + T.colb / 2);
                    if constexpr (NT)
                        __builtin_nontemporal_store(hc | (lc << 8), cp_);
                    else
                        *cp_ = hc | (lc << 8);
                }
            }
            bu = mn2(bu, 1.0);
            bd = mn2(bd, 1.0);
        };
        double ax_k = 0, ay_k = 0, az_k = 0;
        double bu_m = 0, bd_m = 0, axl_m = 0, ayd_m = 0;
        if (doA && (fxk || fyk)) {
            faces(L, T, wkr, wku, fxk, fyk, r1, ax_k, ay_k, az_k);
            if (fxk) sm.ax[ka][L] = ax_k;
            if (fyk) sm.ay[ka][L] = ay_k;
            if constexpr (SAVE && TILE)
                if (store && k >= z0 && k < z1) {
                    put<NT>(a.Ax + zm1.off, T.colb, ax_k);
                    put<NT>(a.Ay + zm1.off, T.colb, ay_k);
                    put<NT>(a.Az + zm1.off, T.colb, az_k);
                }
        }
        if (doB && r1) {
            ratios(L, T, bu_m, bd_m, axl_m, ayd_m);
            sm.bu[ma][L] = bu_m;
            sm.bd[ma][L] = bd_m;
        }
        // limited donor pass of plane o on the tile
        if constexpr (TILE) {
            if (doO) {
                const int ob4 = (p8 - 3) & 3, oa = (p8 - 3) & 1;
                const double* bu = sm.bu[oa];
                const double* bd = sm.bd[oa];
                const double* co = sm.c1[ob4];
                const double b0u = T.bu_m3, b0d = T.bd_m3;
                const double uxr = limit_face_clamped(T.ax_m3, b0u, b0d, bu[L + 1], bd[L + 1]);
                const double uxl = limit_face_clamped(T.axl_m3, bu[L - 1], bd[L - 1], b0u, b0d);
                const double vyr = limit_face_clamped(T.ay_m3, b0u, b0d, bu[L + RX], bd[L + RX]);
                const double vyl = limit_face_clamped(T.ayd_m3, bu[L - RX], bd[L - RX], b0u, b0d);
                const double wzr = limit_face_clamped(T.az_m3, b0u, b0d, bu_m, bd_m);
                const double wzl = limit_face_clamped(T.az_m4, T.bu_m4, T.bd_m4, b0u, b0d);
                double r = donor(T.c1_m3, co[L - 1], co[L + 1], co[L - RX], co[L + RX], T.c1_m4,
                                 T.c1_m2, uxl, uxr, vyl, vyr, wzl, wzr);
                if (store) {
                    if (EMIT && a.q) r = r + qo;
                    put<NT>(a.cout + zm3.off, T.colb, r);
                    if (a.Lx) {
                        put<NT>(a.Lx + zm3.off, T.colb, uxr);
                        put<NT>(a.Ly + zm3.off, T.colb, vyr);
                        put<NT>(a.Lz + zm3.off, T.colb, wzr);
                    }
                }
            }
        }
        // rotate the per-column registers
        T.cn_m3 = T.cn_m2;
        T.cn_m2 = T.cn_m1;
        T.cn_m1 = T.cn_0;
        T.cn_0 = T.cn_p1;
        T.w_m2 = T.w_m1;
        T.w_m1 = T.w_0;
        T.wr_m = wkr;
        T.wu_m = wku;
        T.u_m1 = T.u_0;
        T.v_m1 = T.v_0;
        T.c1_m4 = T.c1_m3;
        T.c1_m3 = T.c1_m2;
        T.c1_m2 = T.c1_m1;
        T.c1_m1 = T.c1_0;
        T.ax_m3 = T.ax_m2;
        T.ay_m3 = T.ay_m2;
        T.ax_m2 = ax_k;
        T.ay_m2 = ay_k;
        T.az_m4 = T.az_m3;
        T.az_m3 = T.az_m2;
        T.az_m2 = az_k;
        if constexpr (TILE) {
            T.axl_m3 = axl_m;
            T.ayd_m3 = ayd_m;
            T.bu_m4 = T.bu_m3;
            T.bd_m4 = T.bd_m3;
            T.bu_m3 = bu_m;
            T.bd_m3 = bd_m;
        }
        zp2.next();
        zp1.next();
        zm2.next();
        zm3.next();
        zp0.next();
        zm1.next();
    };
    using Steady = std::integral_constant<bool, true>;
    using Edge = std::integral_constant<bool, false>;
    int pl = pstart;
    const int s0 = z0 + 3, s1 = z1 + 1;  // steady planes
    for (; pl < s0 && pl <= pend; pl++) plane(pl, Edge());
    for (; pl + 3 <= s1; pl += 4) {
        plane(pl, Steady());
        plane(pl + 1, Steady());
        plane(pl + 2, Steady());
        plane(pl + 3, Steady());
    }
    for (; pl <= s1; pl++) plane(pl, Steady());
    for (; pl <= pend; pl++) plane(pl, Edge());
}

// Tile waves first, then the ring waves; both roles take the same planes and barriers.
// EMIT: the pass adds the emission increment (a.q set; the last pass of a step).
template <int TX, int TY, bool DONOR, bool NT, bool SAVE, bool EMIT>
__global__ void __launch_bounds__((Geo3<TX, TY>::NTH)) k_f12(FusedArgs a, Work w)
{
    using G = Geo3<TX, TY>;
    __shared__ F12Lds<G::R> sm;
    const bool tile = __builtin_amdgcn_readfirstlane(threadIdx.x) < G::NT;
    for (int it = blockIdx.x; it < w.nitems; it += gridDim.x) {
        int item = item_of(w, it);
        FusedArgs x = a;
        if (w.ninst) {
            ConstInt* item0 = (ConstInt*)w.item0;
            int i = find_inst(item0, w.ninst, item);
            __builtin_memcpy(&x, (const FusedArgs*)((ConstArgs*)w.table + i), sizeof x);
            item -= item0[i];
        }
        if (tile)
            f12_part<TX, TY, DONOR, NT, SAVE, EMIT, true>(x, item, threadIdx.x, sm);
        else
            f12_part<TX, TY, DONOR, NT, SAVE, EMIT, false>(x, item, threadIdx.x - G::NT, sm);
        lds_barrier();
    }
}

// ---------------------------------------------------------------- dispatch ---

// Launches a kernel; with ev set, between the start event ev[0] and the stop event
// ev[1], which then bracket the kernel's execution alone.
template <typename K, typename... Args>
void launch_k(K kernel, dim3 grid, dim3 block, hipStream_t s, const hipEvent_t* ev,
              Args... args)
{
    if (ev)
        hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev[0], ev[1], 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

template <int TX, int TY, bool NT>
void launch_nt(const FusedArgs& a, const Work& w, int grid, bool f1, bool donor, hipStream_t s,
               const hipEvent_t* ev)
{
    using G = Geo<TX, TY>;
    if (f1) {
        if (donor)
            launch_k(k_f1<TX, TY, true, NT>, dim3(grid), dim3(G::NT), s, ev, a, w);
        else
            launch_k(k_f1<TX, TY, false, NT>, dim3(grid), dim3(G::NT), s, ev, a, w);
    } else if (a.bu) {
        launch_k(k_f2<TX, TY, true, NT>, dim3(grid), dim3(G::NT), s, ev, a, w);
    } else {
        launch_k(k_f2<TX, TY, false, NT>, dim3(grid), dim3(G::NT), s, ev, a, w);
    }
}

template <int TX, int TY>
int launch_pair(const FusedArgs& a, const Work& w, int grid, bool f1, bool donor, hipStream_t s,
                const hipEvent_t* ev)
{
    if (a.nt)
        launch_nt<TX, TY, true>(a, w, grid, f1, donor, s, ev);
    else
        launch_nt<TX, TY, false>(a, w, grid, f1, donor, s, ev);
    MPD_HIP(hipGetLastError());
    return MPD_OK;
}

int launch_any(const FusedArgs& a, const Work& w, int grid, int tx, int ty, bool f1, bool donor,
               hipStream_t s, const hipEvent_t* ev = nullptr)
{
#define MPD_TILE(X, Y) \
    if (tx == X && ty == Y) return launch_pair<X, Y>(a, w, grid, f1, donor, s, ev);
    MPD_FUSED_TILES(MPD_TILE)
#undef MPD_TILE
    return fail(MPD_ERR_ARG, "tile %dx%d is not one of the compiled tile shapes", tx, ty);
}

// Tile shapes of the one-kernel-per-pass path (LDS: 23 planes of the tile + 3-cell ring).
constexpr int F12_TX = F12_TILE_X, F12_TY = F12_TILE_Y;

template <bool DONOR, bool SAVE>
void launch_f12_as(const FusedArgs& a, const Work& w, int grid, bool emit, hipStream_t s,
                   const hipEvent_t* ev)
{
    constexpr int TX = F12_TX, TY = F12_TY;
    const dim3 blk(Geo3<TX, TY>::NTH);
    if (a.nt) {
        if (emit)
            launch_k(k_f12<TX, TY, DONOR, true, SAVE, true>, dim3(grid), blk, s, ev, a, w);
        else
            launch_k(k_f12<TX, TY, DONOR, true, SAVE, false>, dim3(grid), blk, s, ev, a, w);
    } else {
        if (emit)
            launch_k(k_f12<TX, TY, DONOR, false, SAVE, true>, dim3(grid), blk, s, ev, a, w);
        else
            launch_k(k_f12<TX, TY, DONOR, false, SAVE, false>, dim3(grid), blk, s, ev, a, w);
    }
}

// emit: the pass adds an emission increment (for a batch: some instance's descriptor has
// one; the others' are null)
int launch_f12(const FusedArgs& a, const Work& w, int grid, bool donor, bool emit, hipStream_t s,
               const hipEvent_t* ev)
{
    if (a.bu) {  // the adjoint's record (first pass only)
        if (!donor) return fail(MPD_ERR_UNSUPPORTED, "records of the first pass only");
        launch_f12_as<true, true>(a, w, grid, emit, s, ev);
    } else if (donor) {
        launch_f12_as<true, false>(a, w, grid, emit, s, ev);
    } else {
        launch_f12_as<false, false>(a, w, grid, emit, s, ev);
    }
    MPD_HIP(hipGetLastError());
    return MPD_OK;
}

// single grid: one work item per workgroup
int launch(FusedArgs& a, int tx, int ty, bool f1, bool donor, hipStream_t s, const hipEvent_t* ev)
{
    fused_geometry(a, tx, ty);
    const int nb = a.tiles_x * a.tiles_y * a.nchunks;
    Work w{nullptr, nullptr, 0, nb, nb % 8 == 0 ? 1 : 0};
    return launch_any(a, w, nb, tx, ty, f1, donor, s, ev);
}

// The start/stop events of the next launch when the grid's step is being timed.
int span(mpd_grid_s* g, hipEvent_t ev[2], const hipEvent_t** use)
{
    *use = nullptr;
    if (!g->trace) return MPD_OK;
    MPD_HIP(hipEventCreate(&ev[0]));
    MPD_HIP(hipEventCreate(&ev[1]));
    g->trace->push_back(ev[0]);
    g->trace->push_back(ev[1]);
    g->trace_pairs = true;
    *use = ev;
    return MPD_OK;
}

// The unlimited corrective donor pass cout = donor(cin; U, V, W) (+ q), in launches of
// at most DMAXZ planes (the timing events, when given, bracket a single launch).
int donor_faces(const mpd_grid_s* g, const double* U, const double* V, const double* W,
                const double* cin, const double* q, double* cout, hipStream_t s,
                const hipEvent_t* ev)
{
    if ((g->ny + DBY - 1) / DBY > 65535)
        return fail(MPD_ERR_UNSUPPORTED, "unlimited periodic passes need ny <= %d", 65535 * DBY);
    const dim3 grid((g->nx + DBX - 1) / DBX, (g->ny + DBY - 1) / DBY, 1);
    if (g->nz <= DMAXZ) {
        launch_k(k_donor_faces, dim3(grid.x, grid.y, g->nz), dim3(DBX, DBY), s, ev, g->nx,
                 g->ny, g->nz, 0, U, V, W, cin, q, cout);
    } else {
        if (ev) MPD_HIP(hipEventRecord(ev[0], s));
        for (int k0 = 0; k0 < g->nz; k0 += DMAXZ)
            launch_k(k_donor_faces, dim3(grid.x, grid.y, std::min(DMAXZ, g->nz - k0)),
                     dim3(DBX, DBY), s, nullptr, g->nx, g->ny, g->nz, k0, U, V, W, cin, q, cout);
        if (ev) MPD_HIP(hipEventRecord(ev[1], s));
    }
    MPD_HIP(hipGetLastError());
    return MPD_OK;
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

// z-chunk of the one-kernel-per-pass kernel.  One workgroup fills a CU's LDS, so the
// items run in rounds of one per CU, and a chunk of zc planes takes zc + 5 plane
// iterations (the pipeline's start and end): the chunk count c is the one that
// minimises rounds(cols * c) * (zc + 5), zc = ceil(nz / c), fewer items on a tie.
// (Halving the whole column until every CU had an item left 192^3 with 288 items, two
// rounds for 1.1 rounds of work: 0.279 ms per step against 0.24 on two kernels per pass.)
int choose_zc_f12(const mpd_grid_s* g)
{
    static int cus = 0;
    if (!cus) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            cus = n;
        else
            cus = 256;
    }
    const long long cols =
        (long long)((g->nx + F12_TX - 1) / F12_TX) * ((g->ny + F12_TY - 1) / F12_TY);
    int best = g->nz;
    long long best_cost = -1;
    for (int c = 1; c <= g->nz; c++) {
        const int zc = (g->nz + c - 1) / c;
        if ((g->nz + zc - 1) / zc != c) continue;  // the same chunks as a smaller count
        const long long rounds = (cols * c + cus - 1) / cus;
        const long long cost = rounds * (zc + 5);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = zc;
        }
    }
    return best;
}

}  // namespace

#ifdef MPD_FAST_DIV
// (one definition: the fast build's)
int pass_zchunk(const mpd_grid_s* g) { return g->zc3 ? g->zc3 : choose_zc_f12(g); }
#endif

int MPD_CAT(fused_batch_launch, MPD_ARITH)(int tx, int ty, bool f1, bool donor,
                                           const FusedArgs* table, const int* item0, int ninst,
                                           int nitems, int grid, bool nt, hipStream_t s)
{
    Work w{table, item0, ninst, nitems, (nitems % 8 == 0 && grid % 8 == 0) ? 1 : 0};
    FusedArgs none{};
    none.nt = nt;
    return launch_any(none, w, grid, tx, ty, f1, donor, s);
}

int MPD_CAT(fused_batch_launch_f12, MPD_ARITH)(const FusedArgs* table, const int* item0,
                                               int ninst, int nitems, bool donor, bool emit,
                                               bool nt, hipStream_t s)
{
    Work w{table, item0, ninst, nitems, nitems % 8 == 0 ? 1 : 0};
    FusedArgs none{};
    none.nt = nt;
    return launch_f12(none, w, nitems, donor, emit, s, nullptr);
}

int MPD_CAT(fused_resident_blocks, MPD_ARITH)(int tx, int ty, bool f1)
{
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
#define MPD_TILE(X, Y)                                                                      \
    if (tx == X && ty == Y)                                                                 \
        e = f1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_f1<X, Y, true, false>,       \
                                                             Geo<X, Y>::NT, 0)              \
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_f2<X, Y, false, false>,      \
                                                             Geo<X, Y>::NT, 0);
    MPD_FUSED_TILES(MPD_TILE)
#undef MPD_TILE
    return e == hipSuccess ? n : 0;
}

int MPD_CAT(fused_store_step, MPD_ARITH)(mpd_grid_s* g, const double* cin, double* cout,
                                         double* c1, double* const A[3], double* bu,
                                         double* bd, int* code)
{
    hipStream_t s = stream();
    int rc;
    const int tx = g->tile_x ? g->tile_x : 32, ty = g->tile_y ? g->tile_y : 8;
    FusedArgs a{};
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.np = (long long)g->nx * g->ny;
    a.zc = g->zc ? g->zc : choose_zc(g, tx, ty);
    a.nt = fused_streaming(g->ncell);
    a.cin = cin;
    a.U = g->face[0].p;
    a.V = g->face[1].p;
    a.W = g->face[2].p;
    a.c1out = c1;
    a.Ax = A[0];
    a.Ay = A[1];
    a.Az = A[2];
    if (one_kernel_passes(g)) {
        // the grid's steps take one kernel per pass: so does its record step
        a.cn = cin;
        a.q = g->has_q ? g->q.p : nullptr;
        a.cout = cout;
        a.bu = bu;
        a.bd = bd;
        a.code = code;
        a.zc = g->zc3 ? g->zc3 : choose_zc_f12(g);
        fused_geometry(a, F12_TX, F12_TY);
        const int nb = fused_items(a);
        const Work w{nullptr, nullptr, 0, nb, nb % 8 == 0 ? 1 : 0};
        return launch_f12(a, w, nb, true, a.q != nullptr, s, nullptr);
    }
    if ((rc = launch(a, tx, ty, true, true, s, nullptr))) return rc;
    const double* q = g->has_q ? g->q.p : nullptr;
    if (g->nonosc) {
        a.cn = cin;
        a.cp = c1;
        a.q = q;
        a.cout = cout;
        a.bu = bu;
        a.bd = bd;
        a.code = code;
        // the limiter kernel in its own launch shape, as in the plain step
        const int tx2 = g->tile2_x ? g->tile2_x : tx, ty2 = g->tile2_x ? g->tile2_y : ty;
        a.zc = g->tile2_x ? (g->zc2 ? g->zc2 : choose_zc(g, tx2, ty2)) : a.zc;
        return launch(a, tx2, ty2, false, false, s, nullptr);
    }
    return donor_faces(g, A[0], A[1], A[2], c1, q, cout, s, nullptr);
}

int MPD_CAT(periodic_fused_step, MPD_ARITH)(mpd_grid_s* g)
{
    hipStream_t s = launch_stream(g);
    int rc;
    if ((rc = ensure_fused_scratch(g))) return rc;
    int tx = g->tile_x ? g->tile_x : 32, ty = g->tile_y ? g->tile_y : 8;
    FusedArgs a{};
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.np = (long long)g->nx * g->ny;
    a.nt = fused_streaming(g->ncell);
    const int zc1 = g->zc ? g->zc : choose_zc(g, tx, ty);
    // F2's launch shape: its own when set, else F1's
    const int tx2 = g->tile2_x ? g->tile2_x : tx, ty2 = g->tile2_x ? g->tile2_y : ty;
    const int zc2 = g->tile2_x ? (g->zc2 ? g->zc2 : choose_zc(g, tx2, ty2)) : zc1;
    const double* q = g->has_q ? g->q.p : nullptr;
    double* cin = g->c[g->cur].p;
    double* out = g->c[1 - g->cur].p;
    const double* prev[3] = {g->face[0].p, g->face[1].p, g->face[2].p};
    const double* field = cin;
    if (one_kernel_passes(g)) {
        // one kernel per corrective pass; the limited faces of pass p (read by pass p+1)
        // alternate between the two face sets
        const int zc = g->zc3 ? g->zc3 : choose_zc_f12(g);
        a.zc = zc;
        fused_geometry(a, F12_TX, F12_TY);
        const int nb = fused_items(a);
        const Work w{nullptr, nullptr, 0, nb, nb % 8 == 0 ? 1 : 0};
        for (int p = 1; p < g->iord; p++) {
            const bool last = (p == g->iord - 1);
            mpd::DevBuf* lf = (p & 1) ? g->fb : g->fa;
            a.cn = cin;
            a.cin = field;
            a.U = prev[0];
            a.V = prev[1];
            a.W = prev[2];
            a.q = last ? q : nullptr;
            a.cout = last ? out : g->cp[p & 1].p;
            a.Lx = last ? nullptr : lf[0].p;
            a.Ly = last ? nullptr : lf[1].p;
            a.Lz = last ? nullptr : lf[2].p;
            hipEvent_t ev[2];
            const hipEvent_t* use;
            if ((rc = span(g, ev, &use))) return rc;
            if ((rc = launch_f12(a, w, nb, p == 1, a.q != nullptr, s, use))) return rc;
            for (int d = 0; d < 3; d++) prev[d] = lf[d].p;
            field = a.cout;
        }
        g->cur = 1 - g->cur;
        return MPD_OK;
    }
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
        a.zc = zc1;
        hipEvent_t ev[2];
        const hipEvent_t* use;
        if ((rc = span(g, ev, &use))) return rc;
        if ((rc = launch(a, tx, ty, true, p == 1, s, use))) return rc;
        const double* pass_in = (p == 1) ? g->cp[0].p : field;
        if ((rc = span(g, ev, &use))) return rc;
        if (g->nonosc) {
            a.cn = cin;
            a.cp = pass_in;
            a.q = last ? q : nullptr;
            a.cout = dst;
            a.Lx = last ? nullptr : g->fb[0].p;
            a.Ly = last ? nullptr : g->fb[1].p;
            a.Lz = last ? nullptr : g->fb[2].p;
            a.zc = zc2;
            if ((rc = launch(a, tx2, ty2, false, false, s, use))) return rc;
            for (int d = 0; d < 3; d++) prev[d] = g->fb[d].p;
        } else {
            if ((rc = donor_faces(g, tgt[0].p, tgt[1].p, tgt[2].p, pass_in,
                                  last ? q : nullptr, dst, s, use)))
                return rc;
            for (int d = 0; d < 3; d++) prev[d] = tgt[d].p;
        }
        field = dst;
    }
    g->cur = 1 - g->cur;
    return MPD_OK;
}

}  // namespace mpd
