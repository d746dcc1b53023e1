// Host-side helpers shared by the library's translation units: error reporting,
// the per-device launch stream and a checked-allocation helper.
#pragma once
#include <new>
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <string>

#include "../../include/mpdata.h"

namespace mpd {

// Records the calling thread's last error message and returns the code.
int fail(int code, const char* fmt, ...);
void clear_error();

// HIP status -> MPD code with a message naming the call site.
int hip_fail(hipError_t e, const char* what);

// Bodies of the C ABI's entry points: a C++ exception (a failed host allocation) never
// crosses into a C caller, it becomes an error code and message.
#define MPD_API_BEGIN try {
#define MPD_API_END                                                                        \
    }                                                                                      \
    catch (const std::bad_alloc&)                                                          \
    {                                                                                      \
        return ::mpd::fail(MPD_ERR_NOMEM, "host memory allocation failed");                \
    }                                                                                      \
    catch (...)                                                                            \
    {                                                                                      \
        return ::mpd::fail(MPD_ERR_DEVICE, "unexpected C++ exception in the library");    \
    }

#define MPD_HIP(call)                                                   \
    do {                                                                \
        hipError_t e_ = (call);                                         \
        if (e_ != hipSuccess) return ::mpd::hip_fail(e_, #call);        \
    } while (0)

// Ensures a usable device is current; returns MPD_OK or MPD_ERR_DEVICE.
int ensure_device();
// The stream every kernel of this library is launched on (one per device, default
// priority, created on first use).
hipStream_t stream();

// Device buffer owned by RAII.
struct DevBuf {
    double* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int alloc(size_t count);  // no-op when already count elements
    void release();
};

// Autotuner results: n integers per key ("key v0 v1 ..." lines), cached per process.
// Lookups read the file named by the environment variable MPD_TUNE_CACHE or, when it is
// unset, the table shipped beside the library (tuning.txt); new results are appended to
// the MPD_TUNE_CACHE file only.
bool tune_lookup(const std::string& key, int n, int* v);
void tune_store(const std::string& key, int n, const int* v);

}  // namespace mpd
