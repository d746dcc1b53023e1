// The device-resident grid object behind mpd_grid, smolar() and msmolar().
#pragma once
#include <vector>

#include "mpd_common.hpp"

struct mpd_grid_s {
    int nx = 0, ny = 0, nz = 0;
    int boundary = MPD_PERIODIC;
    int iord = 2;
    int nonosc = 0;
    int path = 0;  // 0 auto, 1 staged, 2 fused (two kernels per pass), 3 one kernel per pass
    int arith = MPD_ARITH_FAST;
    int adj_scatter = 0;  // adjoint: 0 deterministic gathers, 1 atomic scatter
    int adj_tiled = 1;    // adjoint limiter stages: 1 one LDS-tiled kernel, 0 per-cell passes
    long long adj_tape = 0;  // gradient driver: bytes for kept segment records (-1: free memory)
    int adj_kept = 0;        // segment records the last gradient run kept from its forward
    int tile_x = 0, tile_y = 0, zc = 0;  // fused tile shape and z-chunk (0: default)
    // tile shape and z-chunk of the limiter kernel F2 (0: those of F1); the two kernels of
    // a pass share no on-chip state, so each runs in the launch shape that suits it
    int tile2_x = 0, tile2_y = 0, zc2 = 0;
    int zc3 = 0;  // z-chunk of the one-kernel-per-pass kernel (0: chosen from the size)
    size_t ncell = 0;
    size_t nface[3] = {0, 0, 0};  // elements per face family (layout of the mode)

    mpd::DevBuf face[3];  // Courant numbers U, V, W
    mpd::DevBuf c[2];     // field ping-pong; c[cur] is the newest
    int cur = 0;
    mpd::DevBuf q;        // emission increment
    bool has_q = false;
    bool has_winds = false, has_field = false;

    // scratch of the staged paths
    mpd::DevBuf fa[3], fb[3];  // antidiffusive velocities, ping-pong
    mpd::DevBuf cp[2];         // intermediate pass fields
    mpd::DevBuf bup, bdn;      // limiter ratios
    mpd::DevBuf red;           // reduction partials

    // step sentinel: every `sentinel_every` steps (0: off) a device check records in
    // *sentinel_first the lowest step count, since the sentinel was armed, after which a
    // value was non-finite (or negative, with sentinel_nonneg); ~0 while none was
    unsigned long long* sentinel_first = nullptr;
    int sentinel_every = 0, sentinel_nonneg = 0;
    long long sentinel_steps = 0;

    // when set, the kernel launches of a step are timed: the fused path launches each
    // kernel between a start and a stop event of its own (pairs appended to *trace,
    // trace_pairs set); the other paths record one event after each launch
    std::vector<hipEvent_t>* trace = nullptr;
    bool trace_pairs = false;

    // time loop captured as a HIP graph (PERIODIC fused path): a block of steps from each
    // field parity, with the launch configuration it was captured with
    struct GraphKey {
        int tx, ty, zc, tx2, ty2, zc2, zc3, arith, path;
        const double* q;
        bool operator==(const GraphKey& o) const
        {
            return tx == o.tx && ty == o.ty && zc == o.zc && tx2 == o.tx2 && ty2 == o.ty2 &&
                   zc2 == o.zc2 && zc3 == o.zc3 && arith == o.arith && path == o.path &&
                   q == o.q;
        }
    };
    // launch stream of the fused step: the library's stream unless a caller supplies its
    // own (the allocation-free entry point launches on the caller's stream)
    bool caller_stream_set = false;
    hipStream_t caller_stream = nullptr;

    bool use_graph = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    GraphKey gkey[2] = {};

    mpd_grid_s() = default;
    mpd_grid_s(const mpd_grid_s&) = delete;
    mpd_grid_s& operator=(const mpd_grid_s&) = delete;
    ~mpd_grid_s()
    {
        for (hipGraphExec_t& e : gexec)
            if (e) (void)hipGraphExecDestroy(e);
        if (sentinel_first) (void)hipFree(sentinel_first);
    }
};

namespace mpd {

int grid_init(mpd_grid_s* g, int nx, int ny, int nz, int boundary, int iord, int nonosc);
// Face Courant numbers from winds already on the device.
int courant_from_winds(mpd_grid_s* g, const double* dx, const double* dy, const double* dz,
                       double dt, const double* u0, const double* u1, const double* v0,
                       const double* v1, const double* w0, const double* w1);
// One step of the grid's boundary mode; advances g->cur.
int frame_step(mpd_grid_s* g);
int periodic_step(mpd_grid_s* g);
int field_stats(mpd_grid_s* g, double* mass, double* cmin, double* cmax);
// Enqueues the step sentinel's check of the current field (step count g->sentinel_steps).
int sentinel_check_launch(mpd_grid_s* g);
// Records a timing event after a kernel launch when g->trace is set.
void trace_mark(mpd_grid_s* g);
// Whether a PERIODIC grid's corrective passes run as one kernel each: path 3, or the
// automatic path on nonoscillatory grids of at least 8 planes (measured on MI355X, iord
// 2, against two kernels per pass: 64^3 0.021 / 0.028 ms, 128^3 0.058 / 0.065, 192^3
// 0.180 / 0.238, 256^3 0.36 / 0.52 ms per step; 512x512x256 iord 3 3.09 / 4.12 ms; but
// a chunk streams five planes beyond its own, and on thin grids that loses: 128x128x1
// 0.017 / 0.010 ms, 512x512x4 0.053 / 0.045, 128x128x8 0.020 / 0.027).  Grids of one
// plane take the kernel's single-plane form instead (plane_kernel_passes).
inline bool one_kernel_passes(const mpd_grid_s* g)
{
    return g->nonosc && (g->path == 3 || (g->path == 0 && g->nz >= 8));
}
// Whether a grid's corrective passes take the single-plane form of that kernel (path 4,
// automatic for nonoscillatory grids of one z-plane).
inline bool plane_kernel_passes(const mpd_grid_s* g)
{
    return g->nonosc && g->nz == 1 && (g->path == 4 || g->path == 0);
}
// The stream a grid's fused step launches on.
inline hipStream_t launch_stream(const mpd_grid_s* g)
{
    return g->caller_stream_set ? g->caller_stream : stream();
}

}  // namespace mpd
