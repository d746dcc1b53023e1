// C ABI of the library: the drop-in smolar()/msmolar()/msmolar_gather() of the
// reference (smolar.h:45-48, 95-99, 132-136) and the device-resident mpd_* interface.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/smolar.h"
#include "mpd_fused.hpp"
#include "mpd_grid.hpp"

namespace mpd {

int periodic_staged_step(mpd_grid_s* g);
int periodic_fused_step_exact(mpd_grid_s* g);
int periodic_fused_step_fast(mpd_grid_s* g);

int grid_init(mpd_grid_s* g, int nx, int ny, int nz, int boundary, int iord, int nonosc)
{
    if (boundary != MPD_FRAME && boundary != MPD_PERIODIC)
        return fail(MPD_ERR_ARG, "boundary must be MPD_FRAME or MPD_PERIODIC");
    if (iord < 1 || iord > 64) return fail(MPD_ERR_ARG, "iord %d outside 1..64", iord);
    if (boundary == MPD_FRAME && (nx < 3 || ny < 3 || nz < 3))
        return fail(MPD_ERR_ARG, "frame grid %dx%dx%d: every dimension must be >= 3", nx, ny,
                    nz);
    if (nx < 1 || ny < 1 || nz < 1)
        return fail(MPD_ERR_ARG, "grid %dx%dx%d: dimensions must be >= 1", nx, ny, nz);
    if ((size_t)nx * ny * nz > ((size_t)1 << 31) - 1)
        return fail(MPD_ERR_ARG, "grid of %zu cells exceeds 2^31-1", (size_t)nx * ny * nz);
    if (nonosc && boundary != MPD_PERIODIC)
        return fail(MPD_ERR_UNSUPPORTED, "the nonoscillatory limiter needs MPD_PERIODIC");
    int rc = ensure_device();
    if (rc) return rc;
    g->nx = nx;
    g->ny = ny;
    g->nz = nz;
    g->boundary = boundary;
    g->iord = iord;
    g->nonosc = nonosc ? 1 : 0;
    g->ncell = (size_t)nx * ny * nz;
    if (boundary == MPD_FRAME) {
        g->nface[0] = (size_t)(nx - 1) * (ny - 2) * (nz - 2);
        g->nface[1] = (size_t)(nx - 2) * (ny - 1) * (nz - 2);
        g->nface[2] = (size_t)(nx - 2) * (ny - 2) * (nz - 1);
    } else {
        g->nface[0] = g->nface[1] = g->nface[2] = g->ncell;
    }
    for (int d = 0; d < 3; d++)
        if ((rc = g->face[d].alloc(g->nface[d])) != MPD_OK) return rc;
    for (int b = 0; b < 2; b++)
        if ((rc = g->c[b].alloc(g->ncell)) != MPD_OK) return rc;
    g->cur = 0;
    return MPD_OK;
}

// Scratch of the fused path: unlimited / limited antidiffusive faces, pass fields.
int ensure_fused_scratch(mpd_grid_s* g)
{
    int rc;
    for (int d = 0; d < 3; d++) {
        if ((rc = g->fa[d].alloc(g->ncell)) != MPD_OK) return rc;
        if ((rc = g->fb[d].alloc(g->ncell)) != MPD_OK) return rc;
    }
    for (int b = 0; b < 2; b++)
        if ((rc = g->cp[b].alloc(g->ncell)) != MPD_OK) return rc;
    return MPD_OK;
}

static int step_once(mpd_grid_s* g)
{
    if (g->boundary == MPD_FRAME) return frame_step(g);
    return periodic_step(g);
}

// The fused kernels serve every corrective pass; a donor-only step (iord 1) is a single
// staged kernel already.
int periodic_step(mpd_grid_s* g)
{
    if (g->path != 1 && g->iord >= 2 && fused_plane_fits(g->nx, g->ny))
        return g->arith == MPD_ARITH_FAST ? periodic_fused_step_fast(g)
                                          : periodic_fused_step_exact(g);
    return periodic_staged_step(g);
}

// Copies between the caller's (pageable) host arrays and the device.  Copies of 8 MiB
// and more go through two pinned staging buffers of 8 MiB: chunks alternate between
// them, the host copying one (on four threads) while the DMA engine moves the other.
// Measured on the MI355X host link with 64 MiB: pageable 28-30 GB/s to the device and
// 17 GB/s back; staged 44 GB/s either way (one host thread: 27-29; pinned DMA alone 56).
// A copy leases a buffer pair from a pool per device (the drop-in smolar() is reentrant
// per thread: concurrent copies each hold a pair of their own); a pair is allocated when
// no free one is left and kept until the process ends, so the pinned memory is bounded
// by the number of copies that ran at the same time, not by the threads that ever
// copied.  h2d returns once the host array has been read (its DMA may still be in
// flight: the pair's next user waits for it); d2h returns with the host array written.
namespace {
// (below 8 MiB the direct pageable copy was as fast: 64^3 smolar() 0.63 ms per call direct,
// 0.78 staged)
constexpr size_t kStageChunk = (size_t)8 << 20, kStageMin = (size_t)8 << 20;
constexpr int kCopyThreads = 4;

// memcpy split over the calling thread and up to kCopyThreads - 1 helpers (a helper that
// cannot be started leaves its part to the calling thread; nothing throws past the ABI)
void host_copy(char* dst, const char* src, size_t len)
{
    if (len < ((size_t)2 << 20)) {
        std::memcpy(dst, src, len);
        return;
    }
    std::thread helper[kCopyThreads - 1];
    bool started[kCopyThreads - 1] = {};
    for (int k = 1; k < kCopyThreads; k++) {
        const size_t a = len * k / kCopyThreads, b = len * (k + 1) / kCopyThreads;
        try {
            helper[k - 1] = std::thread([=] { std::memcpy(dst + a, src + a, b - a); });
            started[k - 1] = true;
        } catch (...) {
            std::memcpy(dst + a, src + a, b - a);
        }
    }
    std::memcpy(dst, src, len / kCopyThreads);
    for (int k = 0; k < kCopyThreads - 1; k++)
        if (started[k]) helper[k].join();
}

struct Staging {
    Staging* next = nullptr;  // free list of the pool
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool busy[2] = {false, false};  // an enqueued DMA may still use the buffer
    int init()
    {
        if (buf[1]) return MPD_OK;
        for (int i = 0; i < 2; i++) {
            if (!buf[i]) MPD_HIP(hipHostMalloc((void**)&buf[i], kStageChunk, hipHostMallocDefault));
            if (!ev[i]) MPD_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
        }
        return MPD_OK;
    }
    // waits for the last DMA of buffer i
    int settle(int i)
    {
        if (busy[i]) MPD_HIP(hipEventSynchronize(ev[i]));
        busy[i] = false;
        return MPD_OK;
    }
    int dma(int i, void* dst, const void* src, size_t len, hipMemcpyKind kind, hipStream_t s)
    {
        MPD_HIP(hipMemcpyAsync(dst, src, len, kind, s));
        MPD_HIP(hipEventRecord(ev[i], s));
        busy[i] = true;
        return MPD_OK;
    }
};
// Free pairs per device (events are recorded on the device's stream, so a pair serves
// one device); released without allocating.
constexpr int kStageDevices = 64;
std::mutex g_stage_mu;
Staging* g_stage_free[kStageDevices] = {};

struct StageLease {
    Staging* s = nullptr;
    int dev = 0;
    // a pair of the current device, buffers allocated; nullptr-safe on failure
    int acquire()
    {
        MPD_HIP(hipGetDevice(&dev));
        if (dev < 0 || dev >= kStageDevices)
            return fail(MPD_ERR_UNSUPPORTED, "staged copy: device %d beyond %d", dev,
                        kStageDevices);
        {
            std::lock_guard<std::mutex> lk(g_stage_mu);
            s = g_stage_free[dev];
            if (s) g_stage_free[dev] = s->next;
        }
        if (!s) s = new Staging();
        return s->init();
    }
    ~StageLease()
    {
        if (!s) return;
        std::lock_guard<std::mutex> lk(g_stage_mu);
        s->next = g_stage_free[dev];
        g_stage_free[dev] = s;
    }
    StageLease() = default;
    StageLease(const StageLease&) = delete;
    StageLease& operator=(const StageLease&) = delete;
};
}  // namespace

static int h2d(double* d, const double* h, size_t n)
{
    const size_t bytes = n * sizeof(double);
    hipStream_t s = stream();
    if (bytes < kStageMin) {
        MPD_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
        return MPD_OK;
    }
    int rc;
    StageLease lease;
    if ((rc = lease.acquire())) return rc;
    Staging& st = *lease.s;
    const char* src = (const char*)h;
    char* dst = (char*)d;
    for (size_t off = 0, k = 0; off < bytes; off += kStageChunk, k++) {
        const int i = (int)(k & 1);
        const size_t len = std::min(kStageChunk, bytes - off);
        if ((rc = st.settle(i))) return rc;
        host_copy(st.buf[i], src + off, len);
        if ((rc = st.dma(i, dst + off, st.buf[i], len, hipMemcpyHostToDevice, s))) return rc;
    }
    return MPD_OK;
}

static int d2h(double* h, const double* d, size_t n)
{
    const size_t bytes = n * sizeof(double);
    hipStream_t s = stream();
    if (bytes < kStageMin) {
        MPD_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
        return MPD_OK;
    }
    int rc;
    StageLease lease;
    if ((rc = lease.acquire())) return rc;
    Staging& st = *lease.s;
    const char* src = (const char*)d;
    char* dst = (char*)h;
    const size_t nch = (bytes + kStageChunk - 1) / kStageChunk;
    auto issue = [&](size_t k) {
        const int i = (int)(k & 1);
        const size_t off = k * kStageChunk, len = std::min(kStageChunk, bytes - off);
        int r = st.settle(i);
        return r ? r : st.dma(i, st.buf[i], src + off, len, hipMemcpyDeviceToHost, s);
    };
    if ((rc = issue(0))) return rc;
    for (size_t k = 0; k < nch; k++) {
        // the next chunk's DMA runs while this chunk is copied out on the host
        if (k + 1 < nch && (rc = issue(k + 1))) return rc;
        const int i = (int)(k & 1);
        if ((rc = st.settle(i))) return rc;
        const size_t off = k * kStageChunk;
        host_copy(dst + off, st.buf[i], std::min(kStageChunk, bytes - off));
    }
    return MPD_OK;
}

static int sync() { MPD_HIP(hipStreamSynchronize(stream())); return MPD_OK; }

// Uploads spacings and the six wind arrays (identical host pointers are uploaded
// once) and computes the face Courant numbers of g.
static int winds_to_courant(mpd_grid_s* g, const double* dx, const double* dy,
                            const double* dz, double dt, const double* const w[6],
                            DevBuf sp[3], DevBuf wd[6])
{
    int rc;
    const int dims[3] = {g->nx, g->ny, g->nz};
    const double* hs[3] = {dx, dy, dz};
    for (int d = 0; d < 3; d++) {
        if ((rc = sp[d].alloc(dims[d])) != MPD_OK) return rc;
        if ((rc = h2d(sp[d].p, hs[d], dims[d])) != MPD_OK) return rc;
    }
    const double* dev[6];
    for (int a = 0; a < 6; a++) {
        int same = -1;
        for (int b = 0; b < a; b++)
            if (w[b] == w[a]) same = b;
        if (same >= 0) {
            dev[a] = dev[same];
            continue;
        }
        if ((rc = wd[a].alloc(g->ncell)) != MPD_OK) return rc;
        if ((rc = h2d(wd[a].p, w[a], g->ncell)) != MPD_OK) return rc;
        dev[a] = wd[a].p;
    }
    rc = courant_from_winds(g, sp[0].p, sp[1].p, sp[2].p, dt, dev[0], dev[1], dev[2], dev[3],
                            dev[4], dev[5]);
    if (rc) return rc;
    g->has_winds = true;
    return MPD_OK;
}

}  // namespace mpd

using namespace mpd;

// ------------------------------------------------------------------ drop-in ABI ---

namespace {
// Device state of smolar(), leased per call from a pool per device: a call takes a free
// state of its grid shape and order when there is one (no allocation: the reference
// mallocs per call, uu.c:115), else any free state, reinitialised, else a new one.
// Concurrent calls from several threads (smolar() is reentrant) hold states of their
// own.  States are kept until the process ends, so their number is bounded by the calls
// that ran at the same time, not by the threads that ever called.
struct SmolarCache {
    SmolarCache* next = nullptr;  // free list of the pool
    mpd_grid_s g;
    DevBuf sp[3], wd[6];
};
constexpr int kCacheDevices = 64;
std::mutex g_cache_mu;
SmolarCache* g_cache_free[kCacheDevices] = {};

struct CacheLease {
    SmolarCache* c = nullptr;
    int dev = 0;
    int acquire(int nx, int ny, int ns, int order)
    {
        MPD_HIP(hipGetDevice(&dev));
        if (dev < 0 || dev >= kCacheDevices)
            return fail(MPD_ERR_UNSUPPORTED, "smolar: device %d beyond %d", dev, kCacheDevices);
        {
            std::lock_guard<std::mutex> lk(g_cache_mu);
            SmolarCache** at = &g_cache_free[dev];
            for (SmolarCache** q = at; *q; q = &(*q)->next) {
                const mpd_grid_s& g = (*q)->g;
                if (g.nx == nx && g.ny == ny && g.nz == ns && g.iord == order) {
                    at = q;
                    break;
                }
            }
            c = *at;
            if (c) *at = c->next;
        }
        if (!c) c = new SmolarCache;
        return MPD_OK;
    }
    ~CacheLease()
    {
        if (!c) return;
        std::lock_guard<std::mutex> lk(g_cache_mu);
        c->next = g_cache_free[dev];
        g_cache_free[dev] = c;
    }
    CacheLease() = default;
    CacheLease(const CacheLease&) = delete;
    CacheLease& operator=(const CacheLease&) = delete;
};
}  // namespace

extern "C" int smolar(int nx, int ny, int ns, double* dx, double* dy, double* ds,
                      const double dt, double* u0, double* u1, double* v0, double* v1,
                      double* w0, double* w1, double** c0, double** c1, int flow_type,
                      int order)
{
    MPD_API_BEGIN
    (void)flow_type;  // never read by the reference (uu.c:111)
    clear_error();
    if (!dx || !dy || !ds || !u0 || !u1 || !v0 || !v1 || !w0 || !w1 || !c0 || !c1 || !*c0 ||
        !*c1)
        return fail(MPD_ERR_ARG, "smolar: NULL array argument");
    if (order < 1) return fail(MPD_ERR_ARG, "smolar: order %d < 1", order);
    CacheLease lease;
    int rc;
    if ((rc = lease.acquire(nx, ny, ns, order))) return rc;
    SmolarCache* cache = lease.c;
    mpd_grid_s& g = cache->g;
    if (g.nx != nx || g.ny != ny || g.nz != ns || g.iord != order || !g.c[0].p) {
        cache->g.~mpd_grid_s();
        new (&cache->g) mpd_grid_s();
        if ((rc = grid_init(&g, nx, ny, ns, MPD_FRAME, order, 0)) != MPD_OK) return rc;
    }
    const double* w[6] = {u0, u1, v0, v1, w0, w1};
    if ((rc = winds_to_courant(&g, dx, dy, ds, dt, w, cache->sp, cache->wd))) return rc;
    double* h0 = *c0;
    double* h1 = *c1;
    if ((rc = h2d(g.c[0].p, h0, g.ncell))) return rc;
    g.cur = 0;
    if ((rc = frame_step(&g))) return rc;
    // device c[0] mirrors the caller's first buffer, c[1] the second (smolar.c:44-58)
    if ((rc = d2h(h1, g.c[1].p, g.ncell))) return rc;
    if (order > 1 && (rc = d2h(h0, g.c[0].p, g.ncell))) return rc;
    if ((rc = sync())) return rc;
    if (order % 2 == 0) {
        *c1 = h0;
        *c0 = h1;
    }
    return MPD_OK;
    MPD_API_END
}

extern "C" int msmolar(void** desc, int nx, int ny, int ns, double* dx, double* dy,
                       double* ds, const double dt, double* u0, double* u1, double* v0,
                       double* v1, double* w0, double* w1, double** c0, double** c1,
                       int flow_type, int order, int sx, int sy, int ss, int mode)
{
    MPD_API_BEGIN
    (void)flow_type;
    (void)sx;
    (void)sy;
    (void)ss;
    (void)mode;
    clear_error();
    if (!desc || !c0 || !c1) return fail(MPD_ERR_ARG, "msmolar: NULL argument");
    int rc;
    if (!*desc) {  // msmolar.c:25-32: the first call captures every array
        if (!dx || !dy || !ds || !u0 || !u1 || !v0 || !v1 || !w0 || !w1 || !*c0)
            return fail(MPD_ERR_ARG, "msmolar: NULL array argument on the first call");
        mpd_grid_s* g = new mpd_grid_s;
        if ((rc = grid_init(g, nx, ny, ns, MPD_FRAME, order, 0)) != MPD_OK) {
            delete g;
            return rc;
        }
        DevBuf sp[3], wd[6];
        const double* w[6] = {u0, u1, v0, v1, w0, w1};
        if ((rc = winds_to_courant(g, dx, dy, ds, dt, w, sp, wd)) ||
            (rc = h2d(g->c[0].p, *c0, g->ncell)) || (rc = sync())) {
            delete g;
            return rc;
        }
        g->has_field = true;
        *desc = g;
    }
    mpd_grid_s* g = (mpd_grid_s*)*desc;
    if (g->nx != nx || g->ny != ny || g->nz != ns || g->iord != order)
        return fail(MPD_ERR_ARG, "msmolar: grid or order differs from the first call");
    if ((rc = frame_step(g))) return rc;
    if (order % 2 == 0) {  // msmolar.c:183-187
        double* t = *c0;
        *c0 = *c1;
        *c1 = t;
    }
    return MPD_OK;
    MPD_API_END
}

extern "C" int msmolar_gather(void** desc, int nx, int ny, int ns, double* dx, double* dy,
                              double* ds, const double dt, double* u0, double* u1,
                              double* v0, double* v1, double* w0, double* w1, double** c0,
                              double** c1, int flow_type, int order, int sx, int sy, int ss,
                              int mode)
{
    MPD_API_BEGIN
    (void)dx; (void)dy; (void)ds; (void)dt; (void)u0; (void)u1; (void)v0; (void)v1;
    (void)w0; (void)w1; (void)flow_type; (void)order; (void)sx; (void)sy; (void)ss;
    (void)mode; (void)nx; (void)ny; (void)ns;
    clear_error();
    if (!desc || !*desc || !c0 || !c1 || !*c0 || !*c1)
        return fail(MPD_ERR_ARG, "msmolar_gather: NULL argument or no state");
    mpd_grid_s* g = (mpd_grid_s*)*desc;
    int rc;
    if ((rc = d2h(*c1, g->c[g->cur].p, g->ncell)) ||
        (rc = d2h(*c0, g->c[1 - g->cur].p, g->ncell)) || (rc = sync()))
        return rc;
    return MPD_OK;
    MPD_API_END
}

extern "C" int msmolar_free(void** desc)
{
    MPD_API_BEGIN
    if (desc && *desc) {
        delete (mpd_grid_s*)*desc;
        *desc = nullptr;
    }
    return MPD_OK;
    MPD_API_END
}

// -------------------------------------------------------------- mpd_* interface ---

#define GRID_OR_FAIL(g)                                                  \
    do {                                                                 \
        clear_error();                                                   \
        if (!(g)) return fail(MPD_ERR_ARG, "%s: NULL grid", __func__);   \
    } while (0)

extern "C" int mpd_create(mpd_grid* out, int nx, int ny, int nz, int boundary, int iord,
                          int nonosc)
{
    MPD_API_BEGIN
    clear_error();
    if (!out) return fail(MPD_ERR_ARG, "mpd_create: NULL handle pointer");
    *out = nullptr;
    mpd_grid_s* g = new mpd_grid_s;
    int rc = grid_init(g, nx, ny, nz, boundary, iord, nonosc);
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_destroy(mpd_grid g)
{
    MPD_API_BEGIN
    if (g) {
        (void)hipStreamSynchronize(stream());
        delete g;
    }
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_set_winds(mpd_grid g, const double* dx, const double* dy, const double* dz,
                             double dt, const double* u0, const double* u1, const double* v0,
                             const double* v1, const double* w0, const double* w1)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!dx || !dy || !dz || !u0 || !u1 || !v0 || !v1 || !w0 || !w1)
        return fail(MPD_ERR_ARG, "mpd_set_winds: NULL array");
    DevBuf sp[3], wd[6];
    const double* w[6] = {u0, u1, v0, v1, w0, w1};
    int rc = winds_to_courant(g, dx, dy, dz, dt, w, sp, wd);
    if (rc) return rc;
    return sync();  // the temporaries are freed on return
    MPD_API_END
}

extern "C" int mpd_set_courant(mpd_grid g, const double* U, const double* V, const double* W)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!U || !V || !W) return fail(MPD_ERR_ARG, "mpd_set_courant: NULL array");
    const double* h[3] = {U, V, W};
    int rc;
    for (int d = 0; d < 3; d++)
        if ((rc = h2d(g->face[d].p, h[d], g->nface[d]))) return rc;
    g->has_winds = true;
    return sync();
    MPD_API_END
}

extern "C" int mpd_get_courant(mpd_grid g, double* U, double* V, double* W)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!U || !V || !W) return fail(MPD_ERR_ARG, "mpd_get_courant: NULL array");
    double* h[3] = {U, V, W};
    int rc;
    for (int d = 0; d < 3; d++)
        if ((rc = d2h(h[d], g->face[d].p, g->nface[d]))) return rc;
    return sync();
    MPD_API_END
}

extern "C" int mpd_set_field(mpd_grid g, const double* c)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!c) return fail(MPD_ERR_ARG, "mpd_set_field: NULL array");
    int rc = h2d(g->c[g->cur].p, c, g->ncell);
    if (rc) return rc;
    g->has_field = true;
    return sync();
    MPD_API_END
}

extern "C" int mpd_get_field(mpd_grid g, double* c)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!c) return fail(MPD_ERR_ARG, "mpd_get_field: NULL array");
    int rc = d2h(c, g->c[g->cur].p, g->ncell);
    if (rc) return rc;
    return sync();
    MPD_API_END
}

extern "C" int mpd_get_previous_field(mpd_grid g, double* c)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!c) return fail(MPD_ERR_ARG, "mpd_get_previous_field: NULL array");
    int rc = d2h(c, g->c[1 - g->cur].p, g->ncell);
    if (rc) return rc;
    return sync();
    MPD_API_END
}

extern "C" int mpd_set_emission(mpd_grid g, const double* q)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!q) {
        g->has_q = false;
        return MPD_OK;
    }
    if (g->boundary != MPD_PERIODIC)
        return fail(MPD_ERR_UNSUPPORTED, "emission needs MPD_PERIODIC");
    int rc = g->q.alloc(g->ncell);
    if (rc) return rc;
    if ((rc = h2d(g->q.p, q, g->ncell))) return rc;
    g->has_q = true;
    return sync();
    MPD_API_END
}

extern "C" int mpd_get_emission(mpd_grid g, double* q)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!q) return fail(MPD_ERR_ARG, "mpd_get_emission: NULL array");
    if (!g->has_q) return fail(MPD_ERR_ARG, "mpd_get_emission: no emission is set");
    int rc = d2h(q, g->q.p, g->ncell);
    if (rc) return rc;
    return sync();
    MPD_API_END
}

// Steps per captured graph: even, so a replay leaves the field parity unchanged.
constexpr int kGraphSteps = 8;

static bool graphable(const mpd_grid_s* g)
{
    return g->use_graph && g->boundary == MPD_PERIODIC && g->iord >= 2 && g->path != 1 &&
           fused_plane_fits(g->nx, g->ny) && !g->trace && !g->sentinel_every;
}

// kGraphSteps steps as one graph launch.  The graph of the current parity is captured on
// first use and again whenever the launch configuration has changed since (tile shape,
// z-chunk, arithmetic, path, emission array); it references the grid's buffers, whose
// contents may change freely between replays.  The steps launch kernels only (no memset
// or copy nodes).
static int graph_steps(mpd_grid_s* g)
{
    const mpd_grid_s::GraphKey key{g->tile_x, g->tile_y, g->zc,    g->tile2_x, g->tile2_y,
                                   g->zc2,    g->zc3,    g->arith, g->path,
                                   g->has_q ? g->q.p : nullptr};
    const int par = g->cur;
    hipStream_t s = stream();
    if (!g->gexec[par] || !(g->gkey[par] == key)) {
        if (g->gexec[par]) {
            (void)hipGraphExecDestroy(g->gexec[par]);
            g->gexec[par] = nullptr;
        }
        int rc = ensure_fused_scratch(g);  // allocations stay outside the capture
        if (rc) return rc;
        MPD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
        for (int i = 0; i < kGraphSteps && rc == MPD_OK; i++) rc = periodic_step(g);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s, &graph);
        g->cur = par;  // the capture ran the host side of an even number of steps
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&g->gexec[par], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g->gexec[par] = nullptr;
            return hip_fail(e, "hipGraphInstantiate");
        }
        g->gkey[par] = key;
    }
    MPD_HIP(hipGraphLaunch(g->gexec[par], s));
    return MPD_OK;
}

extern "C" int mpd_step(mpd_grid g, int nsteps)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (nsteps < 0) return fail(MPD_ERR_ARG, "mpd_step: nsteps %d < 0", nsteps);
    if (!g->has_winds || !g->has_field)
        return fail(MPD_ERR_ARG, "mpd_step: set the winds (or Courant numbers) and the field first");
    int done = 0, rc;
    if (graphable(g))
        for (; done + kGraphSteps <= nsteps; done += kGraphSteps)
            if ((rc = graph_steps(g))) return rc;
    for (; done < nsteps; done++) {
        if ((rc = step_once(g))) return rc;
        if (g->sentinel_every && ++g->sentinel_steps % g->sentinel_every == 0 &&
            (rc = sentinel_check_launch(g)))
            return rc;
    }
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_set_sentinel(mpd_grid g, int every, int nonnegative)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (every < 0) return fail(MPD_ERR_ARG, "mpd_set_sentinel: every %d < 0", every);
    if (every && !g->sentinel_first)
        MPD_HIP(hipMalloc((void**)&g->sentinel_first, sizeof(unsigned long long)));
    if (g->sentinel_first)
        MPD_HIP(hipMemsetAsync(g->sentinel_first, 0xff, sizeof(unsigned long long), stream()));
    g->sentinel_every = every;
    g->sentinel_nonneg = nonnegative ? 1 : 0;
    g->sentinel_steps = 0;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_sentinel_check(mpd_grid g, long long* first_bad_step)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (first_bad_step) *first_bad_step = -1;
    if (!g->sentinel_first) return MPD_OK;
    unsigned long long f = ~0ull;
    MPD_HIP(hipMemcpyAsync(&f, g->sentinel_first, sizeof f, hipMemcpyDeviceToHost, stream()));
    MPD_HIP(hipStreamSynchronize(stream()));
    if (f == ~0ull) return MPD_OK;
    if (first_bad_step) *first_bad_step = (long long)f;
    return fail(MPD_ERR_NUMERIC, "step sentinel: %s value in the field after step %llu",
                g->sentinel_nonneg ? "a non-finite or negative" : "a non-finite", f);
    MPD_API_END
}

extern "C" int mpd_set_graph(mpd_grid g, int on)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    g->use_graph = on != 0;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_synchronize(mpd_grid g)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    return sync();
    MPD_API_END
}

extern "C" int mpd_field_device_ptr(mpd_grid g, double** dptr)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!dptr) return fail(MPD_ERR_ARG, "mpd_field_device_ptr: NULL");
    *dptr = g->c[g->cur].p;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_step_timed(mpd_grid g, int nsteps, double* ms)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!ms) return fail(MPD_ERR_ARG, "mpd_step_timed: NULL");
    hipEvent_t a, b;
    MPD_HIP(hipEventCreate(&a));
    MPD_HIP(hipEventCreate(&b));
    int rc = MPD_OK;
    hipError_t e = hipEventRecord(a, stream());
    if (e == hipSuccess) rc = mpd_step(g, nsteps);
    if (e == hipSuccess && rc == MPD_OK) e = hipEventRecord(b, stream());
    if (e == hipSuccess && rc == MPD_OK) e = hipEventSynchronize(b);
    float f = 0.f;
    if (e == hipSuccess && rc == MPD_OK) e = hipEventElapsedTime(&f, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "mpd_step_timed event");
    *ms = f;
    return MPD_OK;
    MPD_API_END
}

namespace mpd {
void trace_mark(mpd_grid_s* g)
{
    if (!g->trace) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    (void)hipEventRecord(e, stream());
    g->trace->push_back(e);
}
}  // namespace mpd

extern "C" int mpd_step_kernel_times(mpd_grid g, int nsteps, double* ms, int max, int* n)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!ms || !n || nsteps < 1 || max < 1)
        return fail(MPD_ERR_ARG, "mpd_step_kernel_times: bad arguments");
    if (g->boundary != MPD_PERIODIC)
        return fail(MPD_ERR_UNSUPPORTED, "mpd_step_kernel_times: PERIODIC grids only");
    std::vector<hipEvent_t> ev;
    hipEvent_t start;
    MPD_HIP(hipEventCreate(&start));
    MPD_HIP(hipEventRecord(start, stream()));
    g->trace = &ev;
    g->trace_pairs = false;
    int rc = mpd_step(g, nsteps);
    g->trace = nullptr;
    // start/stop pairs bracket each kernel alone; single events after each launch give
    // the time since the previous launch's event (launch gaps included)
    const bool pairs = g->trace_pairs;
    const size_t nk = pairs ? ev.size() / 2 : ev.size();
    if (rc == MPD_OK && nk == 0) rc = fail(MPD_ERR_UNSUPPORTED, "no kernel launches traced");
    if (rc == MPD_OK) {
        hipError_t e = hipEventSynchronize(ev.back());
        if (e != hipSuccess) rc = hip_fail(e, "hipEventSynchronize");
    }
    int per = rc == MPD_OK ? (int)(nk / nsteps) : 0;
    if (rc == MPD_OK && per * nsteps != (int)nk)
        rc = fail(MPD_ERR_UNSUPPORTED, "launch count varies between steps");
    if (rc == MPD_OK) {
        int m = per < max ? per : max;
        for (int i = 0; i < m; i++) ms[i] = 0.0;
        hipEvent_t prev = start;
        for (size_t k = 0; k < nk; k++) {
            float f = 0.f;
            if (pairs) {
                (void)hipEventElapsedTime(&f, ev[2 * k], ev[2 * k + 1]);
            } else {
                (void)hipEventElapsedTime(&f, prev, ev[k]);
                prev = ev[k];
            }
            if ((int)(k % per) < m) ms[k % per] += f / nsteps;
        }
        *n = per;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    (void)hipEventDestroy(start);
    return rc;
    MPD_API_END
}

namespace {
__global__ void k_copy(const double* __restrict__ src, double* __restrict__ dst, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (size_t)gridDim.x * blockDim.x)
        dst[t] = src[t];
}
}  // namespace

extern "C" int mpd_calibration_copy(size_t n, int reps, double* ms)
{
    MPD_API_BEGIN
    clear_error();
    if (n == 0 || reps < 1 || !ms) return fail(MPD_ERR_ARG, "mpd_calibration_copy: bad arguments");
    int rc = ensure_device();
    if (rc) return rc;
    DevBuf a, b;
    if ((rc = a.alloc(n)) || (rc = b.alloc(n))) return rc;
    hipStream_t s = stream();
    MPD_HIP(hipMemsetAsync(a.p, 0, n * sizeof(double), s));
    hipEvent_t e0, e1;
    MPD_HIP(hipEventCreate(&e0));
    MPD_HIP(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, s, a.p, b.p, n);  // warm-up
    MPD_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, s, a.p, b.p, n);
    MPD_HIP(hipEventRecord(e1, s));
    MPD_HIP(hipEventSynchronize(e1));
    float f = 0.f;
    MPD_HIP(hipEventElapsedTime(&f, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = f / reps;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_field_stats(mpd_grid g, double* mass, double* cmin, double* cmax)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    return field_stats(g, mass, cmin, cmax);
    MPD_API_END
}

extern "C" int mpd_set_path(mpd_grid g, int path)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (path < 0 || path > 4) return fail(MPD_ERR_ARG, "mpd_set_path: %d not in 0..4", path);
    g->path = path;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_set_arith(mpd_grid g, int arith)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (arith != MPD_ARITH_EXACT && arith != MPD_ARITH_FAST)
        return fail(MPD_ERR_ARG, "mpd_set_arith: %d is neither exact nor fast", arith);
    g->arith = arith;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_set_tile(mpd_grid g, int tile_x, int tile_y)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (tile_x < 0 || tile_y < 0) return fail(MPD_ERR_ARG, "mpd_set_tile: negative tile");
    g->tile_x = tile_x;
    g->tile_y = tile_y;
    g->tile2_x = g->tile2_y = g->zc2 = 0;  // the limiter kernel follows
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_set_limiter_tile(mpd_grid g, int tile_x, int tile_y, int zc)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (tile_x < 0 || tile_y < 0 || zc < 0 || (tile_x == 0) != (tile_y == 0))
        return fail(MPD_ERR_ARG, "mpd_set_limiter_tile: bad tile or z-chunk");
    g->tile2_x = tile_x;
    g->tile2_y = tile_y;
    g->zc2 = tile_x ? zc : 0;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_get_tiles(mpd_grid g, int* v)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!v) return fail(MPD_ERR_ARG, "mpd_get_tiles: NULL array");
    const int t[6] = {g->tile_x, g->tile_y, g->zc, g->tile2_x, g->tile2_y, g->zc2};
    for (int i = 0; i < 6; i++) v[i] = t[i];
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_set_zchunk(mpd_grid g, int zc)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (zc < 0) return fail(MPD_ERR_ARG, "mpd_set_zchunk: negative z-chunk length");
    g->zc = zc;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_set_pass_zchunk(mpd_grid g, int zc)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (zc < 0) return fail(MPD_ERR_ARG, "mpd_set_pass_zchunk: negative z-chunk length");
    g->zc3 = zc;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_get_pass_zchunk(mpd_grid g, int* zc)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!zc) return fail(MPD_ERR_ARG, "mpd_get_pass_zchunk: NULL");
    *zc = pass_zchunk(g);
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_get_path(mpd_grid g, int* path)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (!path) return fail(MPD_ERR_ARG, "mpd_get_path: NULL");
    if (g->boundary == MPD_FRAME)
        *path = g->path == 1 || g->iord < 2 ? 1 : 2;
    else if (g->path == 1 || g->iord < 2 || !fused_plane_fits(g->nx, g->ny))
        *path = 1;
    else
        *path = plane_kernel_passes(g) ? 4 : one_kernel_passes(g) ? 3 : 2;
    return MPD_OK;
    MPD_API_END
}

// Autotuner of one grid: times every compiled tile shape x z-chunk length on the grid
// itself (one warm-up step, then nsteps steps with each kernel between events of its own)
// and keeps, for the antidiffusive kernel F1 and for the limiter kernel F2 separately,
// the launch shape in which that kernel ran fastest: they share no on-chip state, and
// F1 (fewer registers) fills a CU with more workgroups than F2.  It restores both field
// buffers and their order, so tuning leaves the grid unchanged.  Results are cached by
// the grid's shape and options (see tune_lookup): "grid2" entries hold both shapes; a
// "grid" entry (one shape) serves both kernels.
extern "C" int mpd_autotune(mpd_grid g, int nsteps, int* tile_x, int* tile_y, int* zc)
{
    MPD_API_BEGIN
    GRID_OR_FAIL(g);
    if (nsteps < 1) return fail(MPD_ERR_ARG, "mpd_autotune: nsteps < 1");
    if (g->boundary != MPD_PERIODIC || g->iord < 2 || g->path == 1 ||
        !fused_plane_fits(g->nx, g->ny))
        return fail(MPD_ERR_UNSUPPORTED, "mpd_autotune: tunes the fused PERIODIC path only");
    if (!g->has_winds || !g->has_field)
        return fail(MPD_ERR_ARG, "mpd_autotune: set the winds (or Courant numbers) and the field first");
    if (one_kernel_passes(g) || plane_kernel_passes(g)) {
        // the one-kernel pass has a single tile shape and a z-chunk chosen from the size:
        // nothing to time (its trial steps would misread as the two-kernel pair)
        if (tile_x) *tile_x = g->tile_x;
        if (tile_y) *tile_y = g->tile_y;
        if (zc) *zc = g->zc;
        return MPD_OK;
    }
    const std::string opts = "i" + std::to_string(g->iord) + "n" + std::to_string(g->nonosc) +
                             "a" + std::to_string(g->arith) + "q" + std::to_string(g->has_q) +
                             ":" + std::to_string(g->nx) + "x" + std::to_string(g->ny) + "x" +
                             std::to_string(g->nz);
    const std::string key = "grid2:" + opts;
    int rc;
    int c[6];
    if (!tune_lookup(key, 6, c)) {
        if (tune_lookup("grid:" + opts, 3, c)) {
            c[3] = c[4] = c[5] = 0;
        } else {
            hipStream_t s = stream();
            DevBuf save[2];
            for (int b = 0; b < 2; b++) {
                if ((rc = save[b].alloc(g->ncell))) return rc;
                MPD_HIP(hipMemcpyAsync(save[b].p, g->c[b].p, g->ncell * 8, hipMemcpyDeviceToDevice, s));
            }
            const int cur = g->cur;
            // trial steps neither capture graphs nor count for the step sentinel (restored
            // on every way out)
            struct Quiet {
                mpd_grid_s* g;
                bool graph;
                int every;
                long long steps;
                ~Quiet()
                {
                    g->use_graph = graph;
                    g->sentinel_every = every;
                    g->sentinel_steps = steps;
                }
            } quiet{g, g->use_graph, g->sentinel_every, g->sentinel_steps};
            g->use_graph = false;
            g->sentinel_every = 0;
            const int shapes[][2] = {
#define MPD_TILE(X, Y) {X, Y},
                MPD_FUSED_TILES(MPD_TILE)
#undef MPD_TILE
            };
            // per kernel: the two fastest shapes (its kernels of all passes summed); and
            // the fastest shape for both kernels together
            double best[2][2] = {{1e300, 1e300}, {1e300, 1e300}}, bestj = 1e300;
            int bestc[2][2][3], bestjc[3] = {g->tile_x, g->tile_y, g->zc};
            for (int f = 0; f < 2; f++)
                for (int r = 0; r < 2; r++) {
                    bestc[f][r][0] = g->tile_x;
                    bestc[f][r][1] = g->tile_y;
                    bestc[f][r][2] = g->zc;
                }
            for (auto& sh : shapes) {
                int last = 0;
                for (int z : {8, 11, 16, 22, 32, 43, 64}) {
                    z = z < g->nz ? z : g->nz;
                    if (z == last) continue;
                    last = z;
                    g->tile_x = sh[0];
                    g->tile_y = sh[1];
                    g->zc = z;
                    g->tile2_x = g->tile2_y = g->zc2 = 0;
                    // two kernels per corrective pass
                    std::vector<double> ms(2 * (size_t)g->iord, 0.0);
                    int nk = 0;
                    if ((rc = mpd_step(g, 1))) return rc;
                    if ((rc = mpd_step_kernel_times(g, nsteps, ms.data(), (int)ms.size(), &nk)))
                        return rc;
                    nk = std::min(nk, (int)ms.size());
                    double t[2] = {0.0, 0.0};  // kernels alternate F1, F2 (or donor) per pass
                    for (int k = 0; k < nk; k++) t[k & 1] += ms[k];
                    const int here[3] = {sh[0], sh[1], z};
                    for (int f = 0; f < 2; f++) {
                        if (t[f] < best[f][0]) {
                            best[f][1] = best[f][0];
                            std::copy(bestc[f][0], bestc[f][0] + 3, bestc[f][1]);
                            best[f][0] = t[f];
                            std::copy(here, here + 3, bestc[f][0]);
                        } else if (t[f] < best[f][1]) {
                            best[f][1] = t[f];
                            std::copy(here, here + 3, bestc[f][1]);
                        }
                    }
                    if (t[0] + t[1] < bestj) {
                        bestj = t[0] + t[1];
                        std::copy(here, here + 3, bestjc);
                    }
                }
            }
            // Kernel times alone do not rank whole steps reliably (how one kernel leaves
            // the caches for the next depends on both shapes): the leading pairings are
            // timed as whole steps, interleaved over three rounds, best of each kept.
            std::vector<std::array<int, 6>> cand;
            auto add = [&](const int* k1, const int* k2) {
                std::array<int, 6> x = {k1[0], k1[1], k1[2], k2[0], k2[1], k2[2]};
                if (!g->nonosc) x[3] = x[4] = x[5] = 0;  // the second kernel is no F2
                if (std::find(cand.begin(), cand.end(), x) == cand.end()) cand.push_back(x);
            };
            add(bestc[0][0], bestc[1][0]);
            add(bestc[0][0], bestc[1][1]);
            add(bestc[0][1], bestc[1][0]);
            add(bestjc, bestjc);
            const double est = std::max(best[0][0] + best[1][0], 1e-3);
            const int ns = std::max(nsteps, std::min(256, (int)std::ceil(2.0 / est)));
            std::vector<double> tstep(cand.size(), 1e300);
            for (int round = 0; round < 3; round++)
                for (size_t i = 0; i < cand.size(); i++) {
                    g->tile_x = cand[i][0];
                    g->tile_y = cand[i][1];
                    g->zc = cand[i][2];
                    g->tile2_x = cand[i][3];
                    g->tile2_y = cand[i][4];
                    g->zc2 = cand[i][5];
                    double msr = 0.0;
                    if ((rc = mpd_step(g, 1)) || (rc = mpd_step_timed(g, ns, &msr))) return rc;
                    tstep[i] = std::min(tstep[i], msr);
                }
            const size_t win = std::min_element(tstep.begin(), tstep.end()) - tstep.begin();
            for (int b = 0; b < 2; b++)
                MPD_HIP(hipMemcpyAsync(g->c[b].p, save[b].p, g->ncell * 8, hipMemcpyDeviceToDevice, s));
            g->cur = cur;
            MPD_HIP(hipStreamSynchronize(s));
            std::copy(cand[win].begin(), cand[win].end(), c);
        }
        tune_store(key, 6, c);
    }
    g->tile_x = c[0];
    g->tile_y = c[1];
    g->zc = c[2];
    g->tile2_x = c[3];
    g->tile2_y = c[4];
    g->zc2 = c[5];
    if (tile_x) *tile_x = c[0];
    if (tile_y) *tile_y = c[1];
    if (zc) *zc = c[2];
    return MPD_OK;
    MPD_API_END
}
