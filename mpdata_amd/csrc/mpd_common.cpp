#include "mpd_common.hpp"

#include <cstdarg>
#include <mutex>
#include <vector>

namespace mpd {

static thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

void clear_error() { g_err.clear(); }

int hip_fail(hipError_t e, const char* what)
{
    return fail(e == hipErrorOutOfMemory ? MPD_ERR_NOMEM : MPD_ERR_DEVICE, "%s failed: %s",
                what, hipGetErrorString(e));
}

int ensure_device()
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(MPD_ERR_DEVICE, "no HIP device available (%s)",
                    e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    return MPD_OK;
}

static std::mutex g_stream_mu;
static std::vector<hipStream_t> g_streams;

hipStream_t stream()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    std::lock_guard<std::mutex> lk(g_stream_mu);
    if ((int)g_streams.size() <= dev) g_streams.resize(dev + 1, nullptr);
    if (!g_streams[dev]) {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
        g_streams[dev] = s;
    }
    return g_streams[dev];
}

int DevBuf::alloc(size_t count)
{
    if (p && n == count) return MPD_OK;
    release();
    if (count == 0) return MPD_OK;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(double));
    if (e != hipSuccess) {
        p = nullptr;
        return fail(MPD_ERR_NOMEM, "hipMalloc of %zu doubles failed: %s", count,
                    hipGetErrorString(e));
    }
    n = count;
    return MPD_OK;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
}

}  // namespace mpd

extern "C" const char* mpd_error_string(void) { return mpd::g_err.c_str(); }

extern "C" int mpd_device_count(int* n)
{
    if (!n) return mpd::fail(MPD_ERR_ARG, "mpd_device_count: NULL");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return MPD_OK;
}
