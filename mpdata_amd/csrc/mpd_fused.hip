// PERIODIC mode, fused path: two LDS-tiled kernels per corrective pass, and the step
// drivers of both fused paths (the one-kernel-per-pass kernel is in mpd_onepass.hip).
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
#include "mpd_fused_dev.hpp"
#include "mpd_grid.hpp"
#include "mpd_tile.hpp"

#define MPD_CAT2(a, b) a##_##b
#define MPD_CAT(a, b) MPD_CAT2(a, b)

namespace mpd {
int ensure_fused_scratch(mpd_grid_s* g);
using namespace MPD_ARITH;
namespace {



using namespace tile;


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
                const double D = (uk[L + 1] - uk[L - 1]) + (v1 - v2) + (w1 - w2);
                put<NT>(a.Ax + zm1.off, colb,
                        antidiff_s(D, av, cr, c, v1 + v2, b1 - b2, b1 + b2, w1 + w2, e1 - e2,
                                   e1 + e2));
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
                const double D = (vk[L + RX] - vk[L - RX]) + (w1 - w2) + (u1 - (ulu + ul));
                put<NT>(a.Ay + zm1.off, colb,
                        antidiff_s(D, av, cu, c, u1 + (uru + ur), b1 - b2, b1 + b2, w1 + w2,
                                   e1 - e2, e1 + e2));
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
                const double D = (wn - wmm) + (u1 - u2) + (v1 - v2);
                put<NT>(a.Az + zm1.off, colb,
                        antidiff_s(D, av, cf, c, u1 + u2, b1 - b2, b1 + b2, v1 + v2, e1 - e2,
                                   e1 + e2));
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
            const double fxl = F2(nbp[1], c, axl), fxr = F2(c, nbp[2], axr);
            const double fyl = F2(nbp[3], c, ayl), fyr = F2(c, nbp[4], ayr);
            const double fzl = F2(cplo, c, azlo), fzr = F2(c, cphi, azhi);
            beta2(hi, lo, c, fxl, fxr, fyl, fyr, fzl, fzr, bu, bd);
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

// ---------------------------------------------------------------- dispatch ---


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

}  // namespace


int MPD_CAT(fused_batch_launch, MPD_ARITH)(int tx, int ty, bool f1, bool donor,
                                           const FusedArgs* table, const int* item0, int ninst,
                                           int nitems, int grid, bool nt, hipStream_t s)
{
    Work w{table, item0, ninst, nitems, (nitems % 8 == 0 && grid % 8 == 0) ? 1 : 0};
    FusedArgs none{};
    none.nt = nt;
    return launch_any(none, w, grid, tx, ty, f1, donor, s);
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
        a.zc = pass_zchunk(g);
        fused_geometry(a, F12_TILE_X, F12_TILE_Y);
        const int nb = fused_items(a);
        const Work w{nullptr, nullptr, 0, nb, nb % 8 == 0 ? 1 : 0};
        return MPD_CAT(onepass_launch, MPD_ARITH)(a, w, nb, true, a.q != nullptr, s, nullptr);
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
    const bool plane = plane_kernel_passes(g);
    if (plane || one_kernel_passes(g)) {
        // one kernel per corrective pass (on a grid of one plane its single-plane form);
        // the limited faces of pass p (read by pass p+1) alternate between the two face
        // sets
        a.zc = plane ? 1 : pass_zchunk(g);
        fused_geometry(a, F12_TILE_X, F12_TILE_Y);
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
            rc = plane ? MPD_CAT(plane_launch, MPD_ARITH)(a, p == 1, a.q != nullptr, s, use)
                       : MPD_CAT(onepass_launch, MPD_ARITH)(a, w, nb, p == 1, a.q != nullptr,
                                                            s, use);
            if (rc) return rc;
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
