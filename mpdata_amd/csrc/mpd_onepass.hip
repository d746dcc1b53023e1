// PERIODIC mode, one kernel per corrective pass (k_f12): the donor field, the
// antidiffusive velocities, the limiter ratios and the limited donor pass of a pass in
// one LDS-tiled kernel, so that c1 and the unlimited faces never leave the chip.
//
// A translation unit of its own so that it can be built with its own scheduling (the
// compiler's ILP-first machine scheduler: the kernel is bound by LDS and FP64 latency at
// three waves per SIMD, and that schedule interleaves the stages of a plane further),
// without changing the schedule of the two-kernel path.  Compiled twice, as
// mpd_fused.hip: MPD_ARITH=exact (bit-identical to the oracle) and MPD_ARITH=fast.
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
using namespace MPD_ARITH;
namespace {

using namespace tile;

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
    double c1_m3 = 0, c1_m2 = 0, c1_m1 = 0, c1_0 = 0;                    // c1(p-3 .. p)
    double ax_m3 = 0, ay_m3 = 0, ax_m2 = 0, ay_m2 = 0;                   // own A_x, A_y
    double az_m3 = 0, az_m2 = 0;                                         // own A_z
    double axl_m3 = 0, ayd_m3 = 0;    // A_x of the x-, A_y of the y- neighbour, plane p-3
    double bu_m3 = 0, bd_m3 = 0;                                         // own ratios
    // z faces formed once per face, as the upper face of a plane, and carried to the
    // plane above as its lower face: the doubled donor fluxes (F2) of the donor field, of
    // the ratios (unlimited faces) and of the limited donor pass
    double fc_z = 0, fr_z = 0, fl_z = 0;
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
    // the donor field's lower z face of the first plane (pstart), from the loads
    if constexpr (DONOR) T.fc_z = F2(T.cn_m1, T.cn_0, T.w_m1);
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
                if constexpr (DONOR) {
                    double fcu;
                    T.c1_0 = donor_zc(T.cn_0, cn[L - 1], cn[L + 1], cn[L - RX], cn[L + RX],
                                      T.cn_p1, up[L - 1], T.u_0, vp[L - RX], T.v_0, T.w_0,
                                      T.fc_z, fcu);
                    T.fc_z = fcu;
                } else {
                    T.c1_0 = T.fld;
                }
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
                const double D = (uk[M + 1] - uk[M - 1]) + (v1 - v2) + (w1 - w2);
                ox = antidiff_s(D, av, cr, c, v1 + v2, b1 - b2, b1 + b2, w1 + w2, e1 - e2,
                                e1 + e2);
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
                const double D = (vk[M + RX] - vk[M - RX]) + (w1 - w2) + (u1 - (ulu + ul));
                oy = antidiff_s(D, av, cu, c, u1 + (uru + ur), b1 - b2, b1 + b2, w1 + w2,
                                e1 - e2, e1 + e2);
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
                const double D = (C.w_0 - C.w_m2) + (u1 - u2) + (v1 - v2);
                oz = antidiff_s(D, av, cf, c, u1 + u2, b1 - b2, b1 + b2, v1 + v2, e1 - e2,
                                e1 + e2);
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
                          double& ayd, double& fzr_) __attribute__((always_inline)) {
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
            const double fxl = F2(xl, c, axl), fxr = F2(c, xr, C.ax_m2);
            const double fyl = F2(yd, c, ayd), fyr = F2(c, yu, C.ay_m2);
            const double fzl = C.fr_z, fzr = F2(c, C.c1_m1, C.az_m2);
            beta2(hi, lo, c, fxl, fxr, fyl, fyr, fzl, fzr, bu, bd);
            fzr_ = fzr;
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
            double fzr;
            ratios(L, T, bu_m, bd_m, axl_m, ayd_m, fzr);
            T.fr_z = fzr;
            sm.bu[ma][L] = bu_m;
            sm.bd[ma][L] = bd_m;
        } else if (!ST && r1 && m == z0 - 2) {
            // the upper z face of plane z0 - 2: the lower face of the first ratio plane
            T.fr_z = F2(T.c1_m2, T.c1_m1, T.az_m2);
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
                double fzu;
                double r = donor_zc(T.c1_m3, co[L - 1], co[L + 1], co[L - RX], co[L + RX],
                                    T.c1_m2, uxl, uxr, vyl, vyr, wzr, T.fl_z, fzu);
                keep_v(r);
                T.fl_z = fzu;
                if (store) {
                    if (EMIT && a.q) r = r + qo;
                    put<NT>(a.cout + zm3.off, T.colb, r);
                    if (a.Lx) {
                        put<NT>(a.Lx + zm3.off, T.colb, uxr);
                        put<NT>(a.Ly + zm3.off, T.colb, vyr);
                        put<NT>(a.Lz + zm3.off, T.colb, wzr);
                    }
                }
            } else if (!ST && o == z0 - 1) {
                // the upper z face of plane z0 - 1: the lower face of the first output
                // plane (its limited velocity and doubled flux)
                const double wz = limit_face_clamped(T.az_m3, T.bu_m3, T.bd_m3, bu_m, bd_m);
                T.fl_z = F2(T.c1_m3, T.c1_m2, wz);
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
        T.c1_m3 = T.c1_m2;
        T.c1_m2 = T.c1_m1;
        T.c1_m1 = T.c1_0;
        T.ax_m3 = T.ax_m2;
        T.ay_m3 = T.ay_m2;
        T.ax_m2 = ax_k;
        T.ay_m2 = ay_k;
        T.az_m3 = T.az_m2;
        T.az_m2 = az_k;
        if constexpr (TILE) {
            T.axl_m3 = axl_m;
            T.ayd_m3 = ayd_m;
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

// z-chunk of the one-kernel-per-pass kernel.  One workgroup fills a CU's LDS, so the
// items run in rounds of one per CU, and a chunk of zc planes takes zc + 5 plane
// iterations (the pipeline's start and end): the chunk count c is the one that
// minimises rounds(cols * c) * (zc + 5), zc = ceil(nz / c), fewer items on a tie.
// (Halving the whole column until every CU had an item left 192^3 with 288 items, two
// rounds for 1.1 rounds of work: 0.279 ms per step against 0.24 on two kernels per pass.)
int choose_zc_f12(const mpd_grid_s* g)
{
    // the CU count, queried once (a thread-safe static initialisation)
    static const int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            n > 0)
            return n;
        return 256;
    }();
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

// ---------------------------------------------------------------------------------------
// Grids of one z-plane (nz = 1): every z-neighbour of a cell is the cell itself (the
// periodic wrap of one plane), and every "plane below / above" of the one-kernel pass is
// the same plane.  F12's plane pipeline, which streams five planes beyond a chunk to
// keep c1, A and the ratios of three planes in flight, collapses into five
// barrier-separated stages over the one plane: load c^n, U, V, W (and the pass input) of
// the tile and its three-cell ring; c1 on rings 0-2; the faces; the ratios on rings 0-1;
// the limited donor pass on the tile.  The formulas are F12's, with the z-neighbours of a
// cell read at the cell (so the exact build is the oracle's arithmetic bit for bit).
// One workgroup per tile; the threads loop over the cells of each stage.
template <int TX, int TY>
struct Plane2 {
    using G = Geo3<TX, TY>;
    static constexpr int RX = G::RX, RY = G::RY, R = G::R;
    double cn[R], u[R], v[R], w[R], c1[R], ax[R], ay[R], az[R], bu[R], bd[R];
};

constexpr int P2_NTH = 256;

template <int TX, int TY, bool DONOR, bool EMIT>
__global__ void __launch_bounds__(P2_NTH) k_f12_plane(FusedArgs a)
{
    using P = Plane2<TX, TY>;
    constexpr int RX = P::RX, RY = P::RY, R = P::R;
    __shared__ P sm;
    const int bx = blockIdx.x % a.tiles_x, by = blockIdx.x / a.tiles_x;
    auto col = [&](int rx, int ry) {  // element index of region cell (rx, ry)
        const int gx = wrapi(bx * TX + rx - 3, a.nx), gy = wrapi(by * TY + ry - 3, a.ny);
        return (size_t)gy * a.nx + gx;
    };
    // stage 0: the region's inputs
    for (int L = threadIdx.x; L < R; L += P2_NTH) {
        const size_t o = col(L % RX, L / RX);
        sm.cn[L] = a.cn[o];
        sm.u[L] = a.U[o];
        sm.v[L] = a.V[o];
        sm.w[L] = a.W[o];
        if constexpr (!DONOR) sm.c1[L] = a.cin[o];
    }
    __syncthreads();
    // stage 1: the donor field on rings 0-2 (first pass; later passes read their input)
    if constexpr (DONOR) {
        constexpr int W1 = RX - 2, N1 = W1 * (RY - 2);
        for (int i = threadIdx.x; i < N1; i += P2_NTH) {
            const int L = (1 + i / W1) * RX + 1 + i % W1;
            const double c = sm.cn[L], w = sm.w[L];
            sm.c1[L] = donor(c, sm.cn[L - 1], sm.cn[L + 1], sm.cn[L - RX], sm.cn[L + RX], c, c,
                             sm.u[L - 1], sm.u[L], sm.v[L - RX], sm.v[L], w, w);
        }
        __syncthreads();
    }
    // stage 2: antidiffusive velocities (the operand order of F12's faces): x-faces on
    // rows 2..TY+3 from column 1 (the ring-2 cells left of ring 1) to TX+3, y-faces on
    // columns 2..TX+3 from row 1, z-faces on rings 0-1
    {
        constexpr int W2 = TX + 3, N2 = W2 * (TY + 3);  // columns 1..TX+3, rows 1..TY+3
        const double* cc = sm.c1;
        const double* uk = sm.u;
        const double* vk = sm.v;
        for (int i = threadIdx.x; i < N2; i += P2_NTH) {
            const int rx = 1 + i % W2, ry = 1 + i / W2, M = ry * RX + rx;
            const double c = cc[M];
            if (ry >= 2) {  // x-face (uu.c:81-87)
                const double av = uk[M];
                const double vr = vk[M + 1 - RX], vru = vk[M + 1], vu = vk[M], vcc = vk[M - RX];
                const double wr_ = sm.w[M + 1], wrf = wr_, wf = sm.w[M], wc = wf;
                const double cr = cc[M + 1], cru = cc[M + 1 + RX], cu = cc[M + RX];
                const double cd = cc[M - RX], crd = cc[M + 1 - RX];
                const double crf = cr, cf = c, cb = c, crb = cr;
#ifdef MPD_FAST_DIV
                const double v1 = vr + vcc, v2 = vru + vu, w1 = wr_ + wc, w2 = wrf + wf;
                const double b1 = cru + cu, b2 = cd + crd, e1 = crf + cf, e2 = cb + crb;
                const double D = (uk[M + 1] - uk[M - 1]) + (v1 - v2) + (w1 - w2);
                sm.ax[M] = antidiff_s(D, av, cr, c, v1 + v2, b1 - b2, b1 + b2, w1 + w2, e1 - e2,
                                      e1 + e2);
#else
                const double sv = 0.25 * (vr + vru + vu + vcc);
                const double sw = 0.25 * (wr_ + wrf + wf + wc);
                const double div = 0.25 * av * (uk[M + 1] - uk[M - 1] + vr - vru - vu + vcc +
                                                wr_ - wrf - wf + wc);
                sm.ax[M] = antidiff(div, av, cr, c, sv, cru + cu - cd - crd, cru + cu + cd + crd,
                                    sw, crf + cf - cb - crb, crf + cf + cb + crb);
#endif
            }
            if (rx >= 2) {  // y-face (vv.c:85-92)
                const double av = vk[M];
                const double uu = uk[M - 1 + RX], uru = uk[M + RX], ur = uk[M], ucc = uk[M - 1];
                const double ulu = uk[M - 2 + RX], ul = uk[M - 2];
                const double wu_ = sm.w[M + RX], wuf = wu_, wf = sm.w[M], wc = wf;
                const double cu = cc[M + RX], cru = cc[M + 1 + RX], cr = cc[M + 1];
                const double cl = cc[M - 1], clu = cc[M - 1 + RX];
                const double cuf = cu, cf = c, cb = c, cub = cu;
#ifdef MPD_FAST_DIV
                const double u1 = uu + ucc, w1 = wu_ + wc, w2 = wuf + wf;
                const double b1 = cru + cr, b2 = cl + clu, e1 = cuf + cf, e2 = cb + cub;
                const double D = (vk[M + RX] - vk[M - RX]) + (w1 - w2) + (u1 - (ulu + ul));
                sm.ay[M] = antidiff_s(D, av, cu, c, u1 + (uru + ur), b1 - b2, b1 + b2, w1 + w2,
                                      e1 - e2, e1 + e2);
#else
                const double su = 0.25 * (uu + uru + ur + ucc);
                const double sw = 0.25 * (wu_ + wuf + wf + wc);
                const double div = 0.25 * av * (vk[M + RX] - vk[M - RX] + wu_ - wuf - wf + wc +
                                                uu - ulu - ul + ucc);
                sm.ay[M] = antidiff(div, av, cu, c, su, cru + cr - cl - clu, cru + cr + cl + clu,
                                    sw, cuf + cf - cb - cub, cuf + cf + cb + cub);
#endif
            }
            if (rx >= 2 && ry >= 2) {  // z-face (ww.c:248-255)
                const double av = sm.w[M];
                const double ur = uk[M], urf = ur, uf = uk[M - 1], ucc = uf;
                const double vf = vk[M - RX], vuf = vk[M], vu = vuf, vcc = vf;
                const double cf = c, crf = cc[M + 1], cr = crf, cl = cc[M - 1], clf = cl;
                const double cuf = cc[M + RX], cu = cuf, cd = cc[M - RX], cdf = cd;
#ifdef MPD_FAST_DIV
                const double u1 = uf + ucc, u2 = urf + ur, v1 = vf + vcc, v2 = vuf + vu;
                const double b1 = crf + cr, b2 = cl + clf, e1 = cuf + cu, e2 = cd + cdf;
                const double D = (av - av) + (u1 - u2) + (v1 - v2);
                sm.az[M] = antidiff_s(D, av, cf, c, u1 + u2, b1 - b2, b1 + b2, v1 + v2, e1 - e2,
                                      e1 + e2);
#else
                const double su = 0.25 * (ur + urf + uf + ucc);
                const double sv = 0.25 * (vf + vuf + vu + vcc);
                const double div = 0.25 * av * (av - av + uf - urf - ur + ucc + vf - vuf - vu +
                                                vcc);
                sm.az[M] = antidiff(div, av, cf, c, su, crf + cr - cl - clf, crf + cr + cl + clf,
                                    sv, cuf + cu - cd - cdf, cuf + cu + cd + cdf);
#endif
            }
        }
    }
    __syncthreads();
    // stage 3: limiter ratios on rings 0-1, clamped to at most 1 (F12's ratios)
    {
        constexpr int W3 = TX + 2, N3 = W3 * (TY + 2);
        for (int i = threadIdx.x; i < N3; i += P2_NTH) {
            const int M = (2 + i / W3) * RX + 2 + i % W3;
            const double c = sm.c1[M], cn = sm.cn[M];
            const double xl = sm.c1[M - 1], xr = sm.c1[M + 1], yd = sm.c1[M - RX];
            const double yu = sm.c1[M + RX];
            const double hi =
                mx2(mx2(mx2(mx2(cn, c), mx2(sm.cn[M - 1], xl)),
                        mx2(mx2(sm.cn[M + 1], xr), mx2(sm.cn[M - RX], yd))),
                    mx2(mx2(mx2(sm.cn[M + RX], yu), mx2(cn, c)), mx2(cn, c)));
            const double lo =
                mn2(mn2(mn2(mn2(cn, c), mn2(sm.cn[M - 1], xl)),
                        mn2(mn2(sm.cn[M + 1], xr), mn2(sm.cn[M - RX], yd))),
                    mn2(mn2(mn2(sm.cn[M + RX], yu), mn2(cn, c)), mn2(cn, c)));
            const double az = sm.az[M];
            const double fxl = F2(xl, c, sm.ax[M - 1]), fxr = F2(c, xr, sm.ax[M]);
            const double fyl = F2(yd, c, sm.ay[M - RX]), fyr = F2(c, yu, sm.ay[M]);
            const double fzl = F2(c, c, az), fzr = fzl;
            double bu, bd;
            beta2(hi, lo, c, fxl, fxr, fyl, fyr, fzl, fzr, bu, bd);
            sm.bu[M] = mn2(bu, 1.0);
            sm.bd[M] = mn2(bd, 1.0);
        }
    }
    __syncthreads();
    // stage 4: the limited donor pass on the tile (+ emission, the limited faces)
    for (int i = threadIdx.x; i < TX * TY; i += P2_NTH) {
        const int rx = 3 + i % TX, ry = 3 + i / TX, M = ry * RX + rx;
        if (bx * TX + rx - 3 >= a.nx || by * TY + ry - 3 >= a.ny) continue;
        const double b0u = sm.bu[M], b0d = sm.bd[M];
        const double uxr = limit_face_clamped(sm.ax[M], b0u, b0d, sm.bu[M + 1], sm.bd[M + 1]);
        const double uxl = limit_face_clamped(sm.ax[M - 1], sm.bu[M - 1], sm.bd[M - 1], b0u, b0d);
        const double vyr = limit_face_clamped(sm.ay[M], b0u, b0d, sm.bu[M + RX], sm.bd[M + RX]);
        const double vyl =
            limit_face_clamped(sm.ay[M - RX], sm.bu[M - RX], sm.bd[M - RX], b0u, b0d);
        const double wz = limit_face_clamped(sm.az[M], b0u, b0d, b0u, b0d);
        const double c = sm.c1[M];
        double r = donor(c, sm.c1[M - 1], sm.c1[M + 1], sm.c1[M - RX], sm.c1[M + RX], c, c, uxl,
                         uxr, vyl, vyr, wz, wz);
        const size_t o = col(rx, ry);
        if (EMIT && a.q) r = r + a.q[o];
        a.cout[o] = r;
        if (a.Lx) {
            a.Lx[o] = uxr;
            a.Ly[o] = vyr;
            a.Lz[o] = wz;
        }
    }
}

template <int TX, int TY, bool DONOR>
void launch_plane_as(FusedArgs a, bool emit, hipStream_t s, const hipEvent_t* ev)
{
    a.tiles_x = (a.nx + TX - 1) / TX;
    a.tiles_y = (a.ny + TY - 1) / TY;
    const dim3 grid((unsigned)(a.tiles_x * a.tiles_y)), blk(P2_NTH);
    if (emit)
        launch_k(k_f12_plane<TX, TY, DONOR, true>, grid, blk, s, ev, a);
    else
        launch_k(k_f12_plane<TX, TY, DONOR, false>, grid, blk, s, ev, a);
}

// Tile of the single-plane pass: 16 x 16, or 8 x 8 when the plane has fewer 16 x 16
// tiles than the chip has CUs (its stages are then latency, not throughput: 128 x 128,
// iord 2, measured 3.9 us per step in 8 x 8 tiles against 5.7 in 16 x 16 and 7.8 in
// 32 x 16; 512 x 512: 11.4 in 16 x 16, 13.5 in 8 x 8; 2048 x 2048: 107, 143).
template <bool DONOR>
void launch_plane(const FusedArgs& a, bool emit, hipStream_t s, const hipEvent_t* ev)
{
    const long long tiles16 = (long long)((a.nx + 15) / 16) * ((a.ny + 15) / 16);
    if (tiles16 < 256)
        launch_plane_as<8, 8, DONOR>(a, emit, s, ev);
    else
        launch_plane_as<16, 16, DONOR>(a, emit, s, ev);
}

}  // namespace

#ifdef MPD_FAST_DIV
// (one definition: the fast build's)
int pass_zchunk(const mpd_grid_s* g) { return g->zc3 ? g->zc3 : choose_zc_f12(g); }
#endif

int MPD_CAT(onepass_launch, MPD_ARITH)(const FusedArgs& a, const Work& w, int grid, bool donor,
                                       bool emit, hipStream_t s, const hipEvent_t* ev)
{
    return launch_f12(a, w, grid, donor, emit, s, ev);
}

int MPD_CAT(plane_launch, MPD_ARITH)(FusedArgs a, bool donor, bool emit, hipStream_t s,
                                     const hipEvent_t* ev)
{
    if (donor)
        launch_plane<true>(a, emit, s, ev);
    else
        launch_plane<false>(a, emit, s, ev);
    MPD_HIP(hipGetLastError());
    return MPD_OK;
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

}  // namespace mpd
