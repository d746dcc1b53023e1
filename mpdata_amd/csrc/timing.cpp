// Timing API of the reference (timing/timing.c:7-47) and HIP-event device timers.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <ctime>

#include "../../include/timing.h"
#include "mpd_common.hpp"

// timing/timing.h:3 uses CLOCK_REALTIME.
#define MPD_CLOCK CLOCK_REALTIME

extern "C" void smolar_get_timer_resolution(struct smolar_time_t* val)
{
    struct timespec ts;
    clock_getres(MPD_CLOCK, &ts);
    val->seconds = ts.tv_sec;
    val->nanoseconds = ts.tv_nsec;
}

extern "C" void smolar_get_time(struct smolar_time_t* val)
{
    struct timespec ts;
    clock_gettime(MPD_CLOCK, &ts);
    val->seconds = ts.tv_sec;
    val->nanoseconds = ts.tv_nsec;
}

// Difference with a borrow from the seconds when the nanoseconds underflow.
static void diff(const smolar_time_t* a, const smolar_time_t* b, long* s, long* ns)
{
    *s = b->seconds - a->seconds;
    *ns = b->nanoseconds - a->nanoseconds;
    if (*ns < 0) {
        *s -= 1;
        *ns += 1000000000L;
    }
}

extern "C" double smolar_get_time_diff(struct smolar_time_t* val1, struct smolar_time_t* val2)
{
    long s, ns;
    diff(val1, val2, &s, &ns);
    return 1e-9 * (double)ns + (double)s;
}

extern "C" void smolar_print_time_diff(struct smolar_time_t* val1, struct smolar_time_t* val2)
{
    long s, ns;
    diff(val1, val2, &s, &ns);
    printf("%ld.%09ld", s, ns);
}

struct smolar_device_timer_s {
    hipEvent_t a = nullptr, b = nullptr;
};

extern "C" int smolar_device_timer_create(smolar_device_timer* t)
{
    MPD_API_BEGIN
    if (!t) return mpd::fail(MPD_ERR_ARG, "smolar_device_timer_create: NULL");
    int rc = mpd::ensure_device();
    if (rc) return rc;
    auto* x = new smolar_device_timer_s;
    if (hipEventCreate(&x->a) != hipSuccess || hipEventCreate(&x->b) != hipSuccess) {
        delete x;
        return mpd::fail(MPD_ERR_DEVICE, "hipEventCreate failed");
    }
    *t = x;
    return MPD_OK;
    MPD_API_END
}

extern "C" int smolar_device_timer_start(smolar_device_timer t)
{
    MPD_API_BEGIN
    if (!t) return mpd::fail(MPD_ERR_ARG, "NULL timer");
    MPD_HIP(hipEventRecord(t->a, mpd::stream()));
    return MPD_OK;
    MPD_API_END
}

extern "C" int smolar_device_timer_stop(smolar_device_timer t)
{
    MPD_API_BEGIN
    if (!t) return mpd::fail(MPD_ERR_ARG, "NULL timer");
    MPD_HIP(hipEventRecord(t->b, mpd::stream()));
    return MPD_OK;
    MPD_API_END
}

extern "C" int smolar_device_timer_elapsed_ms(smolar_device_timer t, double* ms)
{
    MPD_API_BEGIN
    if (!t || !ms) return mpd::fail(MPD_ERR_ARG, "NULL timer or output");
    MPD_HIP(hipEventSynchronize(t->b));
    float f = 0.f;
    MPD_HIP(hipEventElapsedTime(&f, t->a, t->b));
    *ms = f;
    return MPD_OK;
    MPD_API_END
}

extern "C" int smolar_device_timer_destroy(smolar_device_timer t)
{
    MPD_API_BEGIN
    if (t) {
        (void)hipEventDestroy(t->a);
        (void)hipEventDestroy(t->b);
        delete t;
    }
    return MPD_OK;
    MPD_API_END
}
