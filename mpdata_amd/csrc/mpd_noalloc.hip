// The entry point that allocates nothing: one PERIODIC MPDATA step of the fused path on
// caller-owned device buffers, for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The fused step's scratch -- the first pass's donor field, the intermediate pass field
// and two sets of antidiffusive face arrays, at most 8 arrays of ncell doubles -- lives
// in one file-scope device array of fixed size, each device holding its own copy until
// the process ends.  Host code never uses the array's name as an address (on the host
// it is only a stand-in for the symbol): its device address is looked up once per
// device with hipGetSymbolAddress and kept per device.  Every call on a device shares
// that device's array; calls that overlap on one device race.  Nothing in the array
// needs to start at zero: every scratch element a step reads was written earlier in the
// same step.
#include <mutex>

#include "mpd_fused.hpp"
#include "mpd_grid.hpp"

namespace mpd {

int periodic_fused_step_fast(mpd_grid_s* g);

namespace {

// 256 MiB: 8 arrays of up to 4 Mi cells (iord >= 3), or 4 arrays of up to 8 Mi cells
// (iord 2: one pass needs the donor field and one set of faces).
constexpr size_t kScratchDoubles = size_t(32) << 20;

__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up on the first call on
// that device (with the device current) and kept.
int scratch_ptr(double** out)
{
    int dev = 0;
    MPD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(MPD_ERR_DEVICE, "device %d beyond the %d this entry point serves", dev,
                    kMaxDevices);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        MPD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(MPD_ERR_DEVICE, "hipGetSymbolAddress gave a null scratch address");
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return MPD_OK;
}

// A grid whose buffers all belong to others (the caller's arrays and the scratch array);
// it forgets them before its buffer objects would free them.
struct Borrowed {
    mpd_grid_s g;
    ~Borrowed()
    {
        DevBuf* all[] = {&g.face[0], &g.face[1], &g.face[2], &g.c[0], &g.c[1], &g.q,
                         &g.fa[0],   &g.fa[1],   &g.fa[2],   &g.fb[0], &g.fb[1], &g.fb[2],
                         &g.cp[0],   &g.cp[1]};
        for (DevBuf* b : all) {
            b->p = nullptr;
            b->n = 0;
        }
    }
};

void borrow(DevBuf& b, const double* p, size_t n)
{
    b.p = const_cast<double*>(p);
    b.n = n;
}

}  // namespace
}  // namespace mpd

using namespace mpd;

extern "C" int mpd_step_periodic_noalloc(int nx, int ny, int nz, int iord, int nonosc,
                                         const double* U, const double* V, const double* W,
                                         const double* q, const double* cin, double* cout,
                                         void* stream)
{
    MPD_API_BEGIN
    clear_error();
    if (nx < 1 || ny < 1 || nz < 1)
        return fail(MPD_ERR_ARG, "mpd_step_periodic_noalloc: grid %dx%dx%d", nx, ny, nz);
    if (iord < 2 || iord > 64)
        return fail(MPD_ERR_ARG, "mpd_step_periodic_noalloc: iord %d outside 2..64", iord);
    if (!U || !V || !W || !cin || !cout)
        return fail(MPD_ERR_ARG, "mpd_step_periodic_noalloc: NULL array");
    if (cin == cout)
        return fail(MPD_ERR_ARG, "mpd_step_periodic_noalloc: cin and cout must differ");
    if (!fused_plane_fits(nx, ny))
        return fail(MPD_ERR_UNSUPPORTED, "mpd_step_periodic_noalloc: z-planes above 2^29 cells");
    const size_t ncell = (size_t)nx * ny * nz;
    const size_t arrays = iord == 2 ? 4 : 8;
    if (ncell > kScratchDoubles / arrays)
        return fail(MPD_ERR_ARG,
                    "mpd_step_periodic_noalloc: %zu cells exceed the %zu this entry point "
                    "holds scratch for at iord %d",
                    ncell, kScratchDoubles / arrays, iord);
    double* s = nullptr;
    int rc = scratch_ptr(&s);
    if (rc) return rc;

    Borrowed b;
    mpd_grid_s& g = b.g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.boundary = MPD_PERIODIC;
    g.iord = iord;
    g.nonosc = nonosc ? 1 : 0;
    g.arith = MPD_ARITH_FAST;
    g.ncell = ncell;
    g.nface[0] = g.nface[1] = g.nface[2] = ncell;
    borrow(g.face[0], U, ncell);
    borrow(g.face[1], V, ncell);
    borrow(g.face[2], W, ncell);
    borrow(g.c[0], cin, ncell);
    borrow(g.c[1], cout, ncell);
    g.cur = 0;
    if (q) {
        borrow(g.q, q, ncell);
        g.has_q = true;
    }
    g.has_winds = g.has_field = true;
    // scratch layout: the arrays an iord-2 step uses first (donor field, one face set)
    borrow(g.cp[0], s, ncell);
    for (int d = 0; d < 3; d++) borrow(g.fa[d], s + (1 + d) * ncell, ncell);
    // iord 2 never touches the intermediate pass field or the second face set; they alias
    // the first ones there (non-null, so that the step's scratch check allocates nothing,
    // and inside the arrays that were checked against the scratch size)
    const size_t hi = arrays == 8 ? 4 : 0;
    borrow(g.cp[1], s + hi * ncell, ncell);
    for (int d = 0; d < 3; d++) borrow(g.fb[d], s + (hi + 1 + d) * ncell, ncell);
    g.caller_stream_set = true;
    g.caller_stream = static_cast<hipStream_t>(stream);
    return periodic_fused_step_fast(&g);
    MPD_API_END
}
