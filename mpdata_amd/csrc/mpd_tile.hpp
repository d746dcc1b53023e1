// Device helpers of the 2.5-D streaming kernels (the fused forward kernels and the
// tiled adjoint): the tile + halo geometry, the LDS-only barrier, the periodic plane
// cursor and 32-bit column addressing.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace mpd {
namespace tile {

template <int TX, int TY>
struct Geo {
    static constexpr int RX = TX + 4, RY = TY + 4, R = RX * RY;
    static constexpr int NT = TX * TY;                  // threads: one per tile cell
    static constexpr int H1 = 2 * (TX + 2) + 2 * TY;    // ring-1 cells
    static constexpr int NH = R - NT;                   // halo cells, rings 1 and 2
    static_assert(NH <= NT, "every thread owns at most one halo cell");
    // Role of this thread in work item `item`: the thread numbering rotated by whole
    // waves from one item to the next.  The ring-1 work falls on the lowest-numbered
    // roles, so without the rotation the same waves of every workgroup -- and the same
    // SIMDs of a CU -- would carry it.
    __device__ static int role(int item)
    {
        static_assert(NT % 64 == 0, "whole waves");
        return (int)((threadIdx.x + 64u * (unsigned)(item % (NT / 64))) % (unsigned)NT);
    }
    // Region cell (rx, ry) of halo cell h: ring 1 first (the rows above and below the
    // tile, then the two side columns), then ring 2 in the same order, so that the
    // threads with ring-1 work are the lowest-numbered and fill as few waves as possible.
    __device__ static void halo(int h, int& rx, int& ry)
    {
        if (h < H1) {
            if (h < 2 * (TX + 2)) {
                ry = h < TX + 2 ? 1 : TY + 2;
                rx = 1 + h % (TX + 2);
            } else {
                const int r = h - 2 * (TX + 2);
                ry = 2 + r / 2;
                rx = (r & 1) ? TX + 2 : 1;
            }
        } else {
            int r = h - H1;
            if (r < 2 * (TX + 4)) {
                ry = r < TX + 4 ? 0 : TY + 3;
                rx = r % (TX + 4);
            } else {
                r -= 2 * (TX + 4);
                ry = 1 + r / 2;
                rx = (r & 1) ? TX + 3 : 0;
            }
        }
    }
};

// Geometry with at most one computing cell per thread, for kernels whose ring-1 work
// is as expensive as their tile work: threads [0, NT) own the tile cells, threads
// [NT, NT + H1) the ring-1 cells, so no wave evaluates the expensive stage twice.  The
// ring-2 cells (loads only) are secondary cells of the spare lanes of the last wave and
// then of the first tile threads.
template <int TX, int TY>
struct GeoB {
    using G = Geo<TX, TY>;
    static constexpr int RX = G::RX, RY = G::RY, R = G::R, NT = G::NT, H1 = G::H1;
    static constexpr int NTH = ((NT + H1 + 63) / 64) * 64;  // threads of a workgroup
    static constexpr int H2 = R - NT - H1;                  // ring-2 cells
    static constexpr int SPARE = NTH - NT - H1;
    static_assert(H2 <= SPARE + NT, "every ring-2 cell has a thread");
    // Role of this thread in work item `item`: the numbering rotated by whole waves
    // from item to item, so that the wave left without a tile cell -- idle in the
    // tile-only stages -- is a different one in the workgroups that share a CU.
    __device__ static int role(int item)
    {
        return (int)((threadIdx.x + 64u * (unsigned)(item % (NTH / 64))) % (unsigned)NTH);
    }
    // own cell of thread t: 0 = tile cell, 1 = ring-1 cell, -1 = none
    __device__ static int own(int t, int& rx, int& ry)
    {
        if (t < NT) {
            rx = 2 + t % TX;
            ry = 2 + t / TX;
            return 0;
        }
        if (t < NT + H1) {
            G::halo(t - NT, rx, ry);
            return 1;
        }
        rx = ry = 0;
        return -1;
    }
    // ring-2 cell of thread t, if any
    __device__ static bool second(int t, int& rx, int& ry)
    {
        const int j = t >= NT + H1 ? t - NT - H1 : SPARE + t;
        if ((t >= NT + H1 || t < H2 - SPARE) && j < H2) {
            G::halo(H1 + j, rx, ry);
            return true;
        }
        rx = ry = 0;
        return false;
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

// Work item of workgroup b of a launch of n workgroups, one item each.  Workgroups go
// to the 8 XCDs round robin (b % 8); each XCD is dealt one contiguous range of the item
// list (the first n % 8 ranges one item longer), so tiles that share halo rows run in
// the same L2.  A bijection of [0, n) for every n.
__device__ __forceinline__ int xcd_item(int b, int n)
{
    const int x = b % 8, j = b / 8, q = n / 8, r = n % 8;
    return x * q + min(x, r) + j;
}

__device__ __forceinline__ int wrapi(int a, int n)
{
    a %= n;
    return a < 0 ? a + n : a;
}

// Cursor over the planes of the periodic z axis: the wrapped plane index and the
// element offset of that plane, advanced one plane at a time without a division.  Both
// are wave-uniform (they depend on the loop counter and the instance only), so they
// live in scalar registers and every global access is a uniform plane base plus the
// thread's 32-bit column offset.
struct ZCursor {
    int z, nz;
    size_t off, np;
    __device__ ZCursor(int zz, int nz_, long long np_) : z(wrapi(zz, nz_)), nz(nz_), np((size_t)np_)
    {
        off = (size_t)z * np;
    }
    __device__ void next()
    {
        if (++z == nz) {
            z = 0;
            off = 0;
        } else {
            off += np;
        }
    }
};

// Marks a 32-bit column offset as redefined at this point (an empty asm statement; it
// emits no instruction).  Called at the top of a plane loop, it keeps the offset's
// zero-extension inside the loop, where the loads see "uniform plane base + 32-bit
// offset" and use the scalar-base addressing form; hoisted out of the loop, the offset
// became a 64-bit vector value and every load and store needed a 64-bit vector add.
__device__ __forceinline__ void keep32(unsigned& off) { asm volatile("" : "+v"(off)); }

// A value formed before a conditional use: without it the compiler sinks the whole
// computation (its LDS reads included) into the branch, after everything else in the
// block, where it cannot overlap the other stages.
__device__ __forceinline__ void keep_v(double& x) { asm volatile("" : "+v"(x)); }

// Element of a plane at byte offset colb of the thread's column (a 32-bit offset: the
// host keeps every plane below 4 GiB), so the access is a scalar base plus a vector
// offset.
__device__ __forceinline__ double ld(const double* plane, unsigned colb)
{
    return *(const double*)((const char*)plane + colb);
}
__device__ __forceinline__ void st(double* plane, unsigned colb, double v)
{
    *(double*)((char*)plane + colb) = v;
}
// The same as a streaming (non-temporal) store: the line is written out without being
// kept in the caches, for outputs far larger than the L2 and the Infinity Cache.
__device__ __forceinline__ void st_nt(double* plane, unsigned colb, double v)
{
    __builtin_nontemporal_store(v, (double*)((char*)plane + colb));
}

// Byte offset of the column of region cell (cx, cy) of tile (bx, by) of an nx x ny plane
// (periodic wrap).
template <int TX, int TY>
__device__ __forceinline__ unsigned column(int nx, int ny, int bx, int by, int cx, int cy)
{
    const int gx = wrapi(bx * TX + cx - 2, nx), gy = wrapi(by * TY + cy - 2, ny);
    return (unsigned)(gy * nx + gx) * 8u;
}

}  // namespace tile
}  // namespace mpd
