// Device helpers and launch plumbing shared by the fused kernels (mpd_fused.hip) and the
// one-kernel-per-pass kernel (mpd_onepass.hip).
#pragma once
#include <hip/hip_ext.h>

#include "mpd_fused.hpp"
#include "mpd_tile.hpp"

namespace mpd {

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

// One launch of the one-kernel-per-pass kernel (mpd_onepass.hip) over the work items w
// of descriptor a (or of w's table), in 32 x 16 tiles: donor, the first corrective pass
// (it forms the donor field); emit, the pass adds an emission increment; a.bu set, the
// storing variant (first pass only).  With ev set, between the events ev[0] and ev[1].
int onepass_launch_exact(const FusedArgs& a, const Work& w, int grid, bool donor, bool emit,
                         hipStream_t s, const hipEvent_t* ev);
int onepass_launch_fast(const FusedArgs& a, const Work& w, int grid, bool donor, bool emit,
                        hipStream_t s, const hipEvent_t* ev);
// One launch of the single-plane form of the one-kernel pass over descriptor a of a grid
// of one z-plane (16 x 16 or, on small planes, 8 x 8 tiles, one workgroup each); donor,
// emit and ev as above.
int plane_launch_exact(FusedArgs a, bool donor, bool emit, hipStream_t s,
                       const hipEvent_t* ev);
int plane_launch_fast(FusedArgs a, bool donor, bool emit, hipStream_t s, const hipEvent_t* ev);

namespace {

// A store of the fused kernels: streaming in the NT instantiations (chosen at launch from
// FusedArgs::nt; a run-time branch on the flag measured no gain at all).
template <bool NT>
__device__ __forceinline__ void put(double* plane, unsigned colb, double v)
{
    if constexpr (NT)
        tile::st_nt(plane, colb, v);
    else
        tile::st(plane, colb, v);
}

// tile (bx, by) and z-chunk bz of work item b of an instance
__device__ __forceinline__ void tile_of(const FusedArgs& a, int b, int& bx, int& by, int& bz)
{
    bx = b % a.tiles_x;
    by = (b / a.tiles_x) % a.tiles_y;
    bz = b / (a.tiles_x * a.tiles_y);
}

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

}  // namespace
}  // namespace mpd
