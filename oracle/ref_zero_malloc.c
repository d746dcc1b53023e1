/* Zero-filling malloc for the reference build (oracle/_ref only; test infrastructure).
 *
 * The reference allocates its pseudo-velocity ping-pong buffers with malloc
 * (core/src/uu.c:115, vv.c:120, ww.c:126) and vv_pass never writes column i=1
 * (core/src/vv.c:81-82), so the next pass reads whatever the allocator returned.
 * Linking the reference objects with -Wl,--wrap=malloc sends those calls here, which
 * pins the unwritten column to 0.0 on every call.  The restatement in cpu_ref.c adopts
 * the same contract. */
#include <stdlib.h>

void* __wrap_malloc(size_t n)
{
    return calloc(1, n ? n : 1);
}
