/* Timing API of the reference (timing/timing.h:9-27), plus device-event timers.
 * Link: -lsmolar_hip. */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

/* timing/timing.h:9-13 */
struct smolar_time_t {
    long seconds;
    long nanoseconds;
};

/* timing/timing.h:16 -- resolution of the host clock. */
void smolar_get_timer_resolution(struct smolar_time_t* val);
/* timing/timing.h:19 -- current host time. */
void smolar_get_time(struct smolar_time_t* val);
/* timing/timing.h:22-23 -- val2 - val1 in seconds. */
double smolar_get_time_diff(struct smolar_time_t* val1, struct smolar_time_t* val2);
/* timing/timing.h:26-27 -- prints val2 - val1 as seconds.nanoseconds. */
void smolar_print_time_diff(struct smolar_time_t* val1, struct smolar_time_t* val2);

/* Device timers (no counterpart in the reference): a pair of HIP events recorded on the
 * stream the library launches on, so they bracket exactly the kernels enqueued between
 * start and stop. */
typedef struct smolar_device_timer_s* smolar_device_timer;
int smolar_device_timer_create(smolar_device_timer* t);
int smolar_device_timer_start(smolar_device_timer t);
int smolar_device_timer_stop(smolar_device_timer t);
/* Waits for the stop event; elapsed milliseconds between start and stop. */
int smolar_device_timer_elapsed_ms(smolar_device_timer t, double* ms);
int smolar_device_timer_destroy(smolar_device_timer t);

#ifdef __cplusplus
}
#endif
