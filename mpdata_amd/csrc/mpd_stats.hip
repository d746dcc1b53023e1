// Field health sentinel: mass, minimum, maximum and a count of non-finite values,
// reduced on the device (one partial per workgroup, finished on the host).
#include <cmath>
#include <vector>

#include "mpd_grid.hpp"

namespace mpd {
namespace {

constexpr int TPB = 256;
constexpr int NBLK = 512;

__global__ void k_stats(const double* __restrict__ c, size_t n, double* __restrict__ part)
{
    __shared__ double s_sum[TPB], s_min[TPB], s_max[TPB], s_bad[TPB];
    double sum = 0.0, lo = INFINITY, hi = -INFINITY, bad = 0.0;
    for (size_t t = (size_t)blockIdx.x * TPB + threadIdx.x; t < n; t += (size_t)gridDim.x * TPB) {
        double v = c[t];
        if (!isfinite(v)) {
            bad += 1.0;
            continue;
        }
        sum += v;
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
    s_sum[threadIdx.x] = sum;
    s_min[threadIdx.x] = lo;
    s_max[threadIdx.x] = hi;
    s_bad[threadIdx.x] = bad;
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
            s_min[threadIdx.x] = fmin(s_min[threadIdx.x], s_min[threadIdx.x + w]);
            s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + w]);
            s_bad[threadIdx.x] += s_bad[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x + 0] = s_sum[0];
        part[4 * blockIdx.x + 1] = s_min[0];
        part[4 * blockIdx.x + 2] = s_max[0];
        part[4 * blockIdx.x + 3] = s_bad[0];
    }
}

// Step sentinel: flags the step count `step` in *first (atomic minimum) if any value of
// c is non-finite, or negative when nonneg is set; one atomic per wave that saw one.
__global__ void k_sentinel(const double* __restrict__ c, size_t n, int nonneg,
                           unsigned long long step, unsigned long long* __restrict__ first)
{
    bool bad = false;
    for (size_t t = (size_t)blockIdx.x * TPB + threadIdx.x; t < n; t += (size_t)gridDim.x * TPB) {
        const double v = c[t];
        bad = bad || !isfinite(v) || (nonneg && v < 0.0);
    }
    const unsigned long long b = __ballot(bad);
    if (b && (int)(threadIdx.x % 64) == __ffsll((long long)b) - 1) atomicMin(first, step);
}

}  // namespace

int sentinel_check_launch(mpd_grid_s* g)
{
    hipLaunchKernelGGL(k_sentinel, dim3(NBLK), dim3(TPB), 0, stream(), g->c[g->cur].p, g->ncell,
                       g->sentinel_nonneg, (unsigned long long)g->sentinel_steps,
                       g->sentinel_first);
    MPD_HIP(hipGetLastError());
    return MPD_OK;
}

int field_stats(mpd_grid_s* g, double* mass, double* cmin, double* cmax)
{
    int rc = g->red.alloc(4 * NBLK);
    if (rc) return rc;
    hipStream_t s = stream();
    hipLaunchKernelGGL(k_stats, dim3(NBLK), dim3(TPB), 0, s, g->c[g->cur].p, g->ncell, g->red.p);
    MPD_HIP(hipGetLastError());
    std::vector<double> h(4 * NBLK);
    MPD_HIP(hipMemcpyAsync(h.data(), g->red.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    MPD_HIP(hipStreamSynchronize(s));
    double sum = 0.0, lo = INFINITY, hi = -INFINITY, bad = 0.0;
    for (int b = 0; b < NBLK; b++) {
        sum += h[4 * b];
        lo = std::fmin(lo, h[4 * b + 1]);
        hi = std::fmax(hi, h[4 * b + 2]);
        bad += h[4 * b + 3];
    }
    if (mass) *mass = sum;
    if (cmin) *cmin = lo;
    if (cmax) *cmax = hi;
    if (bad > 0) return fail(MPD_ERR_NUMERIC, "%.0f non-finite values in the field", bad);
    return MPD_OK;
}

}  // namespace mpd
