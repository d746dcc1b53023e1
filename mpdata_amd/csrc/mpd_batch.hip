// Ensembles of advection instances of uneven sizes (batched extension; no counterpart
// in the reference, whose only multi-instance driver loops over subdomains on the host,
// core/src/msmolar.c:44-180).
//
// Each instance is an ordinary periodic grid.  Three ways to advance them:
//   MPD_BATCH_PER_INSTANCE  every instance runs its own launch sequence (two kernels
//                           per corrective pass per instance): the baseline.
//   MPD_BATCH_FUSED         one launch per kernel stage for the whole ensemble: the
//                           fused kernels read a device table of per-instance
//                           descriptors and a persistent grid of workgroups strides
//                           over every (instance, tile, z-chunk) work item.
//   MPD_BATCH_GRAPH         the MPD_BATCH_FUSED time loop captured once in a hipGraph
//                           and replayed (no per-step host launch work).
// The batched launches run one kernel per corrective pass (k_f12) when the ensemble is
// larger than the Infinity Cache, else the two kernels F1, F2.
// Tile shape, z-chunk and grid size come from the built-in autotuner (timed trials on
// the batch itself), or are set explicitly.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "mpd_fused.hpp"
#include "mpd_grid.hpp"

namespace mpd {
int periodic_step(mpd_grid_s* g);
int ensure_fused_scratch(mpd_grid_s* g);
}  // namespace mpd

using namespace mpd;

struct mpd_batch_s {
    int ninst = 0, iord = 2, nonosc = 1, arith = MPD_ARITH_FAST;
    int tx = 32, ty = 8, zc = 16, persistent = 1;
    int tx2 = 0, ty2 = 0, zc2 = 0;  // the limiter kernel's launch shape (0: F1's)
    int path = 0;  // 0 automatic, 2 two kernels per pass, 3 one kernel per pass
    int zc3 = 0;   // z-chunk of the one-kernel launches (0: kPassZchunk)
    int cur = 0;
    long long cells = 0;
    std::vector<mpd_grid_s*> g;
    // device tables: [parity][pass-1][0 = F1, 1 = F2][instance]; first work item of
    // each instance for F1 (ditem0[0 .. ninst)) and for F2 (ditem0[ninst .. 2 ninst)),
    // whose tilings may differ
    FusedArgs* dtab = nullptr;
    int* ditem0 = nullptr;
    int nitems = 0, nitems2 = 0;
    std::vector<FusedArgs> htab;
    std::vector<int> hitem0;
    // cached graphs of the time loop, one per starting field parity
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int g_steps[2] = {-1, -1};
    long long g_key[2] = {-1, -1};
    bool tuned = false;

    ~mpd_batch_s()
    {
        for (auto& e : gexec)
            if (e) (void)hipGraphExecDestroy(e);
        if (dtab) (void)hipFree(dtab);
        if (ditem0) (void)hipFree(ditem0);
        for (auto* x : g) delete x;
    }
};

namespace {

// Default z-chunk of the one-kernel batched launches (at most an instance's height).  On
// the 192-instance ensemble of 64-128-plane instances: 32 planes 7.82 ms per step, 40
// 7.83, 64 7.98, 96 8.00, 24 8.37, 128 8.75, 16 8.73 -- chunks that divide every
// instance evenly give every item the same length.
constexpr int kPassZchunk = 32;

// Whether the batched launches run one kernel per corrective pass: path 3, or the
// automatic path when the ensemble's arrays exceed the Infinity Cache (2^22 cells), as
// for single grids.  Per-instance launches follow each instance's own path.
bool one_kernel(const mpd_batch_s* b)
{
    return b->nonosc && (b->path == 3 || (b->path == 0 && b->cells >= (1LL << 22)));
}

int table_index(const mpd_batch_s* b, int parity, int pass, int f2, int inst)
{
    return ((parity * (b->iord - 1) + (pass - 1)) * 2 + f2) * b->ninst + inst;
}

// the key of the launch configuration a cached graph was captured with
long long config_key(const mpd_batch_s* b)
{
    long long k = ((b->tx * 131LL + b->ty) * 131 + b->zc) * 4 + b->persistent * 2 + b->arith;
    k = ((k * 131 + b->tx2) * 131 + b->ty2) * 131 + b->zc2;
    return (k * 2 + (one_kernel(b) ? 1 : 0)) * 1031 + b->zc3;
}

// Launch shape of the first kernel (f2 = false) or of the limiter kernel (f2 = true).
void shape_of(const mpd_batch_s* b, bool f2, int& tx, int& ty, int& zc)
{
    const bool own = f2 && b->tx2;
    tx = own ? b->tx2 : b->tx;
    ty = own ? b->ty2 : b->ty;
    zc = own ? b->zc2 : b->zc;
}

// Fills the per-instance descriptors of every pass for both field parities (the buffer
// roles of periodic_fused_step) and uploads them.
int build_tables(mpd_batch_s* b)
{
    int rc;
    const int npass = b->iord - 1;
    b->htab.assign((size_t)2 * npass * 2 * b->ninst, FusedArgs{});
    b->hitem0.assign(2 * (size_t)b->ninst, 0);
    int items = 0, items2 = 0;
    int tx2, ty2, zc2;
    shape_of(b, true, tx2, ty2, zc2);
    const bool one = one_kernel(b);
    for (int i = 0; i < b->ninst; i++) {
        mpd_grid_s* g = b->g[i];
        if ((rc = ensure_fused_scratch(g))) return rc;
        if (one) {
            // one descriptor per pass in the F1 slot: the kernel's own tiling, the buffer
            // roles of the single-grid one-kernel step
            FusedArgs geo{};
            geo.nx = g->nx;
            geo.ny = g->ny;
            geo.nz = g->nz;
            geo.np = (long long)g->nx * g->ny;
            geo.nt = fused_streaming((size_t)b->cells);
            geo.zc = std::min(b->zc3 ? b->zc3 : kPassZchunk, g->nz);
            fused_geometry(geo, F12_TILE_X, F12_TILE_Y);
            b->hitem0[i] = items;
            items += fused_items(geo);
            for (int par = 0; par < 2; par++) {
                const double* cin = g->c[par].p;
                double* out = g->c[1 - par].p;
                const double* prev[3] = {g->face[0].p, g->face[1].p, g->face[2].p};
                const double* field = cin;
                for (int p = 1; p <= npass; p++) {
                    const bool last = (p == npass);
                    mpd::DevBuf* lf = (p & 1) ? g->fb : g->fa;
                    FusedArgs f = geo;
                    f.cn = cin;
                    f.cin = field;
                    f.U = prev[0];
                    f.V = prev[1];
                    f.W = prev[2];
                    f.q = (last && g->has_q) ? g->q.p : nullptr;
                    f.cout = last ? out : g->cp[p & 1].p;
                    f.Lx = last ? nullptr : lf[0].p;
                    f.Ly = last ? nullptr : lf[1].p;
                    f.Lz = last ? nullptr : lf[2].p;
                    b->htab[table_index(b, par, p, 0, i)] = f;
                    for (int d = 0; d < 3; d++) prev[d] = lf[d].p;
                    field = f.cout;
                }
            }
            continue;
        }
        FusedArgs geo{};
        // streaming stores when the whole batch's arrays exceed the caches
        geo.nt = fused_streaming((size_t)b->cells);
        geo.nx = g->nx;
        geo.ny = g->ny;
        geo.nz = g->nz;
        geo.np = (long long)g->nx * g->ny;
        FusedArgs geo2 = geo;
        geo.zc = std::min(b->zc, g->nz);
        fused_geometry(geo, b->tx, b->ty);
        geo2.zc = std::min(zc2, g->nz);
        fused_geometry(geo2, tx2, ty2);
        b->hitem0[i] = items;
        b->hitem0[b->ninst + i] = items2;
        items += fused_items(geo);
        items2 += fused_items(geo2);
        for (int par = 0; par < 2; par++) {
            const double* cin = g->c[par].p;
            double* out = g->c[1 - par].p;
            const double* prev[3] = {g->face[0].p, g->face[1].p, g->face[2].p};
            const double* field = cin;
            for (int p = 1; p <= npass; p++) {
                const bool last = (p == npass);
                double* dst = last ? out : g->cp[p & 1].p;
                FusedArgs f1 = geo;
                f1.cin = field;
                f1.U = prev[0];
                f1.V = prev[1];
                f1.W = prev[2];
                f1.c1out = g->cp[0].p;
                f1.Ax = g->fa[0].p;
                f1.Ay = g->fa[1].p;
                f1.Az = g->fa[2].p;
                FusedArgs f2 = f1;
                f2.tiles_x = geo2.tiles_x;
                f2.tiles_y = geo2.tiles_y;
                f2.zc = geo2.zc;
                f2.nchunks = geo2.nchunks;
                f2.cn = cin;
                f2.cp = (p == 1) ? g->cp[0].p : field;
                f2.q = (last && g->has_q) ? g->q.p : nullptr;
                f2.cout = dst;
                f2.Lx = last ? nullptr : g->fb[0].p;
                f2.Ly = last ? nullptr : g->fb[1].p;
                f2.Lz = last ? nullptr : g->fb[2].p;
                b->htab[table_index(b, par, p, 0, i)] = f1;
                b->htab[table_index(b, par, p, 1, i)] = f2;
                for (int d = 0; d < 3; d++) prev[d] = g->fb[d].p;
                field = dst;
            }
        }
    }
    b->nitems = items;
    b->nitems2 = items2;
    if (!b->dtab) {
        MPD_HIP(hipMalloc((void**)&b->dtab, b->htab.size() * sizeof(FusedArgs)));
        MPD_HIP(hipMalloc((void**)&b->ditem0, 2 * (size_t)b->ninst * sizeof(int)));
    }
    hipStream_t s = stream();
    MPD_HIP(hipMemcpyAsync(b->dtab, b->htab.data(), b->htab.size() * sizeof(FusedArgs),
                           hipMemcpyHostToDevice, s));
    MPD_HIP(hipMemcpyAsync(b->ditem0, b->hitem0.data(), b->hitem0.size() * sizeof(int),
                           hipMemcpyHostToDevice, s));
    MPD_HIP(hipStreamSynchronize(s));  // the host tables are reused on the next call
    return MPD_OK;
}

int cu_count()
{
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 256;
    return n;
}

int grid_for(const mpd_batch_s* b, bool f1)
{
    const int items = f1 ? b->nitems : b->nitems2;
    if (!b->persistent) return items;
    int tx, ty, zc;
    shape_of(b, !f1, tx, ty, zc);
    int per = b->arith == MPD_ARITH_FAST ? fused_resident_blocks_fast(tx, ty, f1)
                                         : fused_resident_blocks_exact(tx, ty, f1);
    long long grid = (long long)std::max(per, 1) * cu_count();
    if (grid >= items) return items;
    grid -= grid % 8;
    return (int)std::max<long long>(grid, 1);
}

// one batched step from parity par: launches only (graph-capturable)
int launch_step(mpd_batch_s* b, int par)
{
    hipStream_t s = stream();
    int rc;
    auto launch = b->arith == MPD_ARITH_FAST ? fused_batch_launch_fast : fused_batch_launch_exact;
    if (one_kernel(b)) {
        auto launch1 = b->arith == MPD_ARITH_FAST ? fused_batch_launch_f12_fast
                                                  : fused_batch_launch_f12_exact;
        const bool nt = fused_streaming((size_t)b->cells);
        bool any_q = false;
        for (int i = 0; i < b->ninst; i++) any_q = any_q || b->g[i]->has_q;
        for (int p = 1; p < b->iord; p++)
            if ((rc = launch1(b->dtab + table_index(b, par, p, 0, 0), b->ditem0, b->ninst,
                              b->nitems, p == 1, any_q && p == b->iord - 1, nt, s)))
                return rc;
        return MPD_OK;
    }
    for (int p = 1; p < b->iord; p++) {
        const FusedArgs* t1 = b->dtab + table_index(b, par, p, 0, 0);
        const FusedArgs* t2 = b->dtab + table_index(b, par, p, 1, 0);
        int tx2, ty2, zc2;
        shape_of(b, true, tx2, ty2, zc2);
        const bool nt = fused_streaming((size_t)b->cells);
        if ((rc = launch(b->tx, b->ty, true, p == 1, t1, b->ditem0, b->ninst, b->nitems,
                         grid_for(b, true), nt, s)))
            return rc;
        if ((rc = launch(tx2, ty2, false, false, t2, b->ditem0 + b->ninst, b->ninst,
                         b->nitems2, grid_for(b, false), nt, s)))
            return rc;
    }
    return MPD_OK;
}

void set_parity(mpd_batch_s* b, int par)
{
    b->cur = par;
    for (auto* g : b->g) g->cur = par;
}

// Captures (without running) the nsteps-step loop from the current parity, unless the
// cached graph for that parity already matches.
int ensure_graph(mpd_batch_s* b, int nsteps)
{
    hipStream_t s = stream();
    const long long key = config_key(b);
    const int par = b->cur;
    if (b->gexec[par] && b->g_steps[par] == nsteps && b->g_key[par] == key) return MPD_OK;
    if (b->gexec[par]) {
        (void)hipGraphExecDestroy(b->gexec[par]);
        b->gexec[par] = nullptr;
    }
    hipGraph_t graph;
    MPD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    int rc = MPD_OK, p = par;
    for (int k = 0; k < nsteps && rc == MPD_OK; k++) {
        rc = launch_step(b, p);
        p = 1 - p;
    }
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&b->gexec[par], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        b->gexec[par] = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    b->g_steps[par] = nsteps;
    b->g_key[par] = key;
    return MPD_OK;
}

int run_graph(mpd_batch_s* b, int nsteps)
{
    int rc = ensure_graph(b, nsteps);
    if (rc) return rc;
    MPD_HIP(hipGraphLaunch(b->gexec[b->cur], stream()));
    if (nsteps & 1) set_parity(b, 1 - b->cur);
    return MPD_OK;
}

int run(mpd_batch_s* b, int nsteps, int mode)
{
    int rc;
    for (auto* g : b->g)
        if (!g->has_winds || !g->has_field)
            return fail(MPD_ERR_ARG, "batch step: every instance needs Courant numbers and a field");
    if (mode == MPD_BATCH_PER_INSTANCE) {
        for (int k = 0; k < nsteps; k++) {
            for (auto* g : b->g)
                if ((rc = periodic_step(g))) return rc;
            b->cur = 1 - b->cur;
        }
        return MPD_OK;
    }
    if (!b->nonosc) return fail(MPD_ERR_UNSUPPORTED, "batched launches need the limiter on");
    if ((rc = build_tables(b))) return rc;
    if (mode == MPD_BATCH_GRAPH) return run_graph(b, nsteps);
    for (int k = 0; k < nsteps; k++) {
        if ((rc = launch_step(b, b->cur))) return rc;
        set_parity(b, 1 - b->cur);
    }
    return MPD_OK;
}

}  // namespace

#define BATCH_OR_FAIL(b)                                                 \
    do {                                                                 \
        clear_error();                                                   \
        if (!(b)) return fail(MPD_ERR_ARG, "%s: NULL batch", __func__);  \
    } while (0)

extern "C" int mpd_batch_create(mpd_batch* out, int ninst, const int* dims, int iord, int nonosc)
{
    MPD_API_BEGIN
    clear_error();
    if (!out || !dims || ninst < 1) return fail(MPD_ERR_ARG, "mpd_batch_create: bad arguments");
    *out = nullptr;
    if (iord < 2) return fail(MPD_ERR_ARG, "mpd_batch_create: iord must be >= 2");
    auto* b = new mpd_batch_s;
    b->ninst = ninst;
    b->iord = iord;
    b->nonosc = nonosc ? 1 : 0;
    for (int i = 0; i < ninst; i++) {
        auto* g = new mpd_grid_s;
        b->g.push_back(g);
        int rc = grid_init(g, dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], MPD_PERIODIC, iord,
                           nonosc);
        if (rc == MPD_OK && !fused_plane_fits(g->nx, g->ny))
            rc = fail(MPD_ERR_ARG, "mpd_batch_create: instance %d has more than 2^29 cells "
                                   "per z-plane", i);
        if (rc == MPD_OK) rc = ensure_fused_scratch(g);
        if (rc) {
            delete b;
            return rc;
        }
        b->cells += (long long)g->ncell;
    }
    *out = b;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_batch_destroy(mpd_batch b)
{
    MPD_API_BEGIN
    if (b) {
        (void)hipStreamSynchronize(stream());
        delete b;
    }
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_batch_instance(mpd_batch b, int i, mpd_grid* g)
{
    MPD_API_BEGIN
    BATCH_OR_FAIL(b);
    if (!g || i < 0 || i >= b->ninst) return fail(MPD_ERR_ARG, "mpd_batch_instance: bad index");
    *g = b->g[i];
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_batch_cells(mpd_batch b, long long* cells)
{
    MPD_API_BEGIN
    BATCH_OR_FAIL(b);
    if (!cells) return fail(MPD_ERR_ARG, "mpd_batch_cells: NULL");
    *cells = b->cells;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_batch_set_arith(mpd_batch b, int arith)
{
    MPD_API_BEGIN
    BATCH_OR_FAIL(b);
    if (arith != MPD_ARITH_EXACT && arith != MPD_ARITH_FAST)
        return fail(MPD_ERR_ARG, "mpd_batch_set_arith: bad value");
    b->arith = arith;
    for (auto* g : b->g) g->arith = arith;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_batch_set_tile(mpd_batch b, int tx, int ty, int zc, int persistent)
{
    MPD_API_BEGIN
    BATCH_OR_FAIL(b);
    bool ok = false;
#define MPD_TILE(X, Y) ok = ok || (tx == X && ty == Y);
    MPD_FUSED_TILES(MPD_TILE)
#undef MPD_TILE
    if (!ok || zc < 1) return fail(MPD_ERR_ARG, "mpd_batch_set_tile: %dx%d zc %d", tx, ty, zc);
    b->tx = tx;
    b->ty = ty;
    b->zc = zc;
    b->persistent = persistent ? 1 : 0;
    b->tx2 = b->ty2 = b->zc2 = 0;  // the limiter kernel follows
    for (auto* g : b->g) {
        g->tile_x = tx;
        g->tile_y = ty;
        g->tile2_x = g->tile2_y = g->zc2 = 0;
    }
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_batch_set_limiter_tile(mpd_batch b, int tx, int ty, int zc)
{
    MPD_API_BEGIN
    BATCH_OR_FAIL(b);
    bool ok = tx == 0 && ty == 0;
#define MPD_TILE(X, Y) ok = ok || (tx == X && ty == Y && zc >= 1);
    MPD_FUSED_TILES(MPD_TILE)
#undef MPD_TILE
    if (!ok) return fail(MPD_ERR_ARG, "mpd_batch_set_limiter_tile: %dx%d zc %d", tx, ty, zc);
    b->tx2 = tx;
    b->ty2 = ty;
    b->zc2 = tx ? zc : 0;
    for (auto* g : b->g) {  // per-instance launches use the same tile shapes
        g->tile2_x = tx;
        g->tile2_y = ty;
        g->zc2 = 0;
    }
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_batch_set_path(mpd_batch b, int path)
{
    MPD_API_BEGIN
    BATCH_OR_FAIL(b);
    if (path != 0 && path != 2 && path != 3)
        return fail(MPD_ERR_ARG, "mpd_batch_set_path: %d is not 0, 2 or 3", path);
    b->path = path;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_batch_get_path(mpd_batch b, int* path)
{
    MPD_API_BEGIN
    BATCH_OR_FAIL(b);
    if (!path) return fail(MPD_ERR_ARG, "mpd_batch_get_path: NULL");
    *path = one_kernel(b) ? 3 : 2;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_batch_set_pass_zchunk(mpd_batch b, int zc)
{
    MPD_API_BEGIN
    BATCH_OR_FAIL(b);
    if (zc < 0) return fail(MPD_ERR_ARG, "mpd_batch_set_pass_zchunk: negative");
    b->zc3 = zc;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_batch_get_tiles(mpd_batch b, int* v)
{
    MPD_API_BEGIN
    BATCH_OR_FAIL(b);
    if (!v) return fail(MPD_ERR_ARG, "mpd_batch_get_tiles: NULL array");
    const int t[7] = {b->tx, b->ty, b->zc, b->persistent, b->tx2, b->ty2, b->zc2};
    for (int i = 0; i < 7; i++) v[i] = t[i];
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_batch_step(mpd_batch b, int nsteps, int mode)
{
    MPD_API_BEGIN
    BATCH_OR_FAIL(b);
    if (nsteps < 0 || mode < MPD_BATCH_PER_INSTANCE || mode > MPD_BATCH_GRAPH)
        return fail(MPD_ERR_ARG, "mpd_batch_step: bad nsteps or mode");
    if (nsteps == 0) return MPD_OK;
    return run(b, nsteps, mode);
    MPD_API_END
}

extern "C" int mpd_batch_step_timed(mpd_batch b, int nsteps, int mode, double* ms)
{
    MPD_API_BEGIN
    BATCH_OR_FAIL(b);
    if (!ms || nsteps < 1) return fail(MPD_ERR_ARG, "mpd_batch_step_timed: bad arguments");
    int rc;
    for (auto* g : b->g)
        if (!g->has_winds || !g->has_field)
            return fail(MPD_ERR_ARG, "batch step: every instance needs Courant numbers and a field");
    // tables and the captured graph are prepared outside the timed region
    if (mode != MPD_BATCH_PER_INSTANCE) {
        if (!b->nonosc) return fail(MPD_ERR_UNSUPPORTED, "batched launches need the limiter on");
        if ((rc = build_tables(b))) return rc;
    }
    if (mode == MPD_BATCH_GRAPH && (rc = ensure_graph(b, nsteps))) return rc;
    hipEvent_t e0, e1;
    MPD_HIP(hipEventCreate(&e0));
    MPD_HIP(hipEventCreate(&e1));
    hipStream_t s = stream();
    MPD_HIP(hipEventRecord(e0, s));
    if (mode == MPD_BATCH_GRAPH) {
        rc = run_graph(b, nsteps);
    } else if (mode == MPD_BATCH_FUSED) {
        rc = MPD_OK;
        for (int k = 0; k < nsteps && rc == MPD_OK; k++) {
            rc = launch_step(b, b->cur);
            set_parity(b, 1 - b->cur);
        }
    } else {
        rc = run(b, nsteps, mode);
    }
    if (rc == MPD_OK) {
        MPD_HIP(hipEventRecord(e1, s));
        MPD_HIP(hipEventSynchronize(e1));
        float f = 0.f;
        MPD_HIP(hipEventElapsedTime(&f, e0, e1));
        *ms = f;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
    MPD_API_END
}

namespace {

// The instances' current fields, copied aside (save holds one buffer per instance) before
// timed trials and put back after.
int save_fields(mpd_batch_s* b, std::vector<DevBuf>& save)
{
    int rc;
    hipStream_t s = stream();
    for (int i = 0; i < b->ninst; i++) {
        mpd_grid_s* g = b->g[i];
        if ((rc = save[i].alloc(g->ncell))) return rc;
        MPD_HIP(hipMemcpyAsync(save[i].p, g->c[g->cur].p, g->ncell * 8, hipMemcpyDeviceToDevice,
                               s));
    }
    return MPD_OK;
}

int restore_fields(mpd_batch_s* b, const std::vector<DevBuf>& save, int start)
{
    hipStream_t s = stream();
    set_parity(b, start);
    for (int i = 0; i < b->ninst; i++) {
        mpd_grid_s* g = b->g[i];
        MPD_HIP(hipMemcpyAsync(g->c[g->cur].p, save[i].p, g->ncell * 8, hipMemcpyDeviceToDevice,
                               s));
    }
    MPD_HIP(hipStreamSynchronize(s));
    return MPD_OK;
}

}  // namespace

// Autotuner: times every compiled tile shape x z-chunk x grid kind on the batch itself
// (MPD_BATCH_FUSED, nsteps steps each after one warm-up step) and keeps the fastest; then,
// with that kept for the first kernel, every shape x z-chunk for the limiter kernel
// alone, and with the best of those kept, every shape x z-chunk for the first kernel
// again.  A batch on one kernel per pass has one tile shape: there the z-chunk of the
// batched launches is timed (16 .. 128 planes, up to the tallest instance).  The fields
// are saved before and restored after, so tuning leaves the ensemble unchanged.  Results
// are cached per process by the ensemble's shape signature; with MPD_TUNE_CACHE set they
// are also appended to / read from that file ("batch2" entries hold both kernels'
// shapes, a "batch" entry serves both kernels, "batch3" entries hold the one-kernel
// z-chunk).
extern "C" int mpd_batch_autotune(mpd_batch b, int nsteps, int* tx, int* ty, int* zc,
                                  int* persistent)
{
    MPD_API_BEGIN
    BATCH_OR_FAIL(b);
    if (nsteps < 1) return fail(MPD_ERR_ARG, "mpd_batch_autotune: nsteps < 1");
    std::string opts = "i" + std::to_string(b->iord) + "n" + std::to_string(b->nonosc) + "a" +
                       std::to_string(b->arith);
    for (auto* g : b->g)
        opts += ":" + std::to_string(g->nx) + "x" + std::to_string(g->ny) + "x" +
                std::to_string(g->nz);
    const std::string key = "batch2:" + opts;
    int rc;
    int c[7];
    // one timed trial of the configuration set on the batch
    auto trial = [&](double& ms) {
        int r = mpd_batch_step(b, 1, MPD_BATCH_FUSED);
        return r ? r : mpd_batch_step_timed(b, nsteps, MPD_BATCH_FUSED, &ms);
    };
    if (one_kernel(b)) {
        int z3 = 0;
        if (!tune_lookup("batch3:" + opts, 1, &z3)) {
            int tallest = 1;
            for (auto* g : b->g) tallest = std::max(tallest, g->nz);
            std::vector<DevBuf> save(b->ninst);
            const int start = b->cur;
            if ((rc = save_fields(b, save))) return rc;
            double best = 1e300, ms = 0;
            for (int z : {16, 24, 32, 48, 64, 96, 128}) {
                if (z > tallest) break;
                b->zc3 = z;
                if ((rc = trial(ms))) return rc;
                if (ms < best) {
                    best = ms;
                    z3 = z;
                }
            }
            if ((rc = restore_fields(b, save, start))) return rc;
            tune_store("batch3:" + opts, 1, &z3);
        }
        b->zc3 = z3;
        if (tx) *tx = F12_TILE_X;
        if (ty) *ty = F12_TILE_Y;
        if (zc) *zc = z3 ? z3 : kPassZchunk;
        if (persistent) *persistent = 0;
        b->tuned = true;
        return MPD_OK;
    }
    if (!tune_lookup(key, 7, c)) {
        if (tune_lookup("batch:" + opts, 4, c)) {
            c[4] = c[5] = c[6] = 0;
        } else {
            std::vector<DevBuf> save(b->ninst);
            const int start = b->cur;
            if ((rc = save_fields(b, save))) return rc;
            const int shapes[][2] = {
#define MPD_TILE(X, Y) {X, Y},
                MPD_FUSED_TILES(MPD_TILE)
#undef MPD_TILE
            };
            double best = 1e300, ms = 0;
            int f1[4] = {b->tx, b->ty, b->zc, b->persistent}, f2[3] = {0, 0, 0};
            for (auto& sh : shapes)  // both kernels in one shape
                for (int z : {8, 16, 32, 64})
                    for (int pers : {0, 1}) {
                        if ((rc = mpd_batch_set_tile(b, sh[0], sh[1], z, pers)) || (rc = trial(ms)))
                            return rc;
                        if (ms < best) {
                            best = ms;
                            f1[0] = sh[0], f1[1] = sh[1], f1[2] = z, f1[3] = pers;
                        }
                    }
            for (auto& sh : shapes)  // the limiter kernel in a shape of its own
                for (int z : {8, 16, 32, 64}) {
                    if ((rc = mpd_batch_set_tile(b, f1[0], f1[1], f1[2], f1[3])) ||
                        (rc = mpd_batch_set_limiter_tile(b, sh[0], sh[1], z)) || (rc = trial(ms)))
                        return rc;
                    if (ms < best) {
                        best = ms;
                        f2[0] = sh[0], f2[1] = sh[1], f2[2] = z;
                    }
                }
            if (f2[0])  // and the first kernel again, with the limiter kernel's shape kept
                for (auto& sh : shapes)
                    for (int z : {8, 16, 32, 64}) {
                        if ((rc = mpd_batch_set_tile(b, sh[0], sh[1], z, f1[3])) ||
                            (rc = mpd_batch_set_limiter_tile(b, f2[0], f2[1], f2[2])) ||
                            (rc = trial(ms)))
                            return rc;
                        if (ms < best) {
                            best = ms;
                            f1[0] = sh[0], f1[1] = sh[1], f1[2] = z;
                        }
                    }
            if ((rc = restore_fields(b, save, start))) return rc;
            for (int i = 0; i < 4; i++) c[i] = f1[i];
            for (int i = 0; i < 3; i++) c[4 + i] = f2[i];
        }
        tune_store(key, 7, c);
    }
    if ((rc = mpd_batch_set_tile(b, c[0], c[1], c[2], c[3]))) return rc;
    if (c[4] && (rc = mpd_batch_set_limiter_tile(b, c[4], c[5], c[6]))) return rc;
    if (tx) *tx = c[0];
    if (ty) *ty = c[1];
    if (zc) *zc = c[2];
    if (persistent) *persistent = c[3];
    b->tuned = true;
    return MPD_OK;
    MPD_API_END
}
