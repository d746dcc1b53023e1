// On-device field generators for the benchmark configurations (no counterpart in the
// reference's core; its drivers build fields on the host, generators/src/*.c).
//
//   mpd_generate_wind   cell-centred wind = solid-body rotation omega*(a x (x - x0))
//                       plus curl(A) of a vector potential made of Fourier modes,
//                       turned into face Courant numbers by the reference formula
//                       (uu.c:132-134 with dx = dt = 1) and scaled so the largest
//                       |face Courant number| equals cmax
//   mpd_generate_blobs  sum of Gaussian blobs (optionally with periodic images), written
//                       to the field or to the emission increment
// Cell centres sit at integer coordinates (i, j, k).
#include <cmath>
#include <vector>

#include "mpd_grid.hpp"

namespace mpd {
int field_stats(mpd_grid_s* g, double* mass, double* cmin, double* cmax);
}

using namespace mpd;

namespace {

constexpr int TPB = 256;
constexpr double TWO_PI = 6.283185307179586476925286766559;

struct WindDev {
    double axis[3], centre[3], omega;
    int nm;
    double k[3][MPD_MAX_MODES][3];  // [component][mode][axis] integer wavenumbers
    double amp[3][MPD_MAX_MODES], phase[3][MPD_MAX_MODES];
};

__global__ void k_gen_wind(int nx, int ny, int nz, WindDev s, double* __restrict__ u,
                           double* __restrict__ v, double* __restrict__ w)
{
    const size_t np = (size_t)nx * ny;
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= np * nz) return;
    const double x = (double)(t % nx), y = (double)((t / nx) % ny), z = (double)(t / np);
    const double rx = x - s.centre[0], ry = y - s.centre[1], rz = z - s.centre[2];
    double uu = s.omega * (s.axis[1] * rz - s.axis[2] * ry);
    double vv = s.omega * (s.axis[2] * rx - s.axis[0] * rz);
    double ww = s.omega * (s.axis[0] * ry - s.axis[1] * rx);
    // dA[d][e] = d A_d / d x_e
    double dA[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const double L[3] = {(double)nx, (double)ny, (double)nz};
    for (int d = 0; d < 3; d++)
        for (int m = 0; m < s.nm; m++) {
            const double* kk = s.k[d][m];
            double th = TWO_PI * (kk[0] * x / L[0] + kk[1] * y / L[1] + kk[2] * z / L[2]) +
                        s.phase[d][m];
            double c = s.amp[d][m] * cos(th);
            for (int e = 0; e < 3; e++) dA[d][e] += c * TWO_PI * kk[e] / L[e];
        }
    uu += dA[2][1] - dA[1][2];
    vv += dA[0][2] - dA[2][0];
    ww += dA[1][0] - dA[0][1];
    u[t] = uu;
    v[t] = vv;
    w[t] = ww;
}

__global__ void k_absmax(const double* __restrict__ a, size_t n, double* __restrict__ part)
{
    __shared__ double s[TPB];
    double m = 0.0;
    for (size_t t = (size_t)blockIdx.x * TPB + threadIdx.x; t < n; t += (size_t)gridDim.x * TPB)
        m = fmax(m, fabs(a[t]));
    s[threadIdx.x] = m;
    __syncthreads();
    for (int h = TPB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void k_scale(double* __restrict__ a, size_t n, double f)
{
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t < n) a[t] *= f;
}

struct BlobDev {
    int nb, periodic;
    double c[MPD_MAX_BLOBS][3], inv2s2[MPD_MAX_BLOBS], amp[MPD_MAX_BLOBS];
};

__global__ void k_gen_blobs(int nx, int ny, int nz, BlobDev b, double* __restrict__ out)
{
    const size_t np = (size_t)nx * ny;
    size_t t = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (t >= np * nz) return;
    const double x = (double)(t % nx), y = (double)((t / nx) % ny), z = (double)(t / np);
    double acc = 0.0;
    const int r = b.periodic ? 1 : 0;
    for (int q = 0; q < b.nb; q++)
        for (int a = -r; a <= r; a++)
            for (int bb = -r; bb <= r; bb++)
                for (int cc = -r; cc <= r; cc++) {
                    double dx = x - b.c[q][0] - a * nx, dy = y - b.c[q][1] - bb * ny;
                    double dz = z - b.c[q][2] - cc * nz;
                    acc += b.amp[q] * exp(-(dx * dx + dy * dy + dz * dz) * b.inv2s2[q]);
                }
    out[t] = acc;
}

unsigned nblk(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }

}  // namespace

extern "C" int mpd_generate_wind(mpd_grid g, const mpd_wind_spec* w)
{
    MPD_API_BEGIN
    clear_error();
    if (!g || !w) return fail(MPD_ERR_ARG, "mpd_generate_wind: NULL argument");
    if (g->boundary != MPD_PERIODIC)
        return fail(MPD_ERR_UNSUPPORTED, "mpd_generate_wind: PERIODIC grids only");
    if (w->nmodes < 0 || w->nmodes > MPD_MAX_MODES || !(w->cmax > 0))
        return fail(MPD_ERR_ARG, "mpd_generate_wind: nmodes 0..%d and cmax > 0", MPD_MAX_MODES);
    WindDev s{};
    for (int d = 0; d < 3; d++) {
        s.axis[d] = w->axis[d];
        s.centre[d] = w->centre[d];
    }
    s.omega = w->omega;
    s.nm = w->nmodes;
    for (int d = 0; d < 3; d++)
        for (int m = 0; m < w->nmodes; m++) {
            for (int e = 0; e < 3; e++) s.k[d][m][e] = w->k[d][m][e];
            s.amp[d][m] = w->amp[d][m];
            s.phase[d][m] = w->phase[d][m];
        }
    int rc;
    DevBuf cw[3], sp[3];
    const int dims[3] = {g->nx, g->ny, g->nz};
    for (int d = 0; d < 3; d++) {
        if ((rc = cw[d].alloc(g->ncell)) || (rc = sp[d].alloc(dims[d]))) return rc;
        std::vector<double> ones(dims[d], 1.0);
        MPD_HIP(hipMemcpy(sp[d].p, ones.data(), dims[d] * sizeof(double), hipMemcpyHostToDevice));
    }
    hipStream_t st = stream();
    hipLaunchKernelGGL(k_gen_wind, dim3(nblk(g->ncell)), dim3(TPB), 0, st, g->nx, g->ny, g->nz, s,
                       cw[0].p, cw[1].p, cw[2].p);
    MPD_HIP(hipGetLastError());
    if ((rc = courant_from_winds(g, sp[0].p, sp[1].p, sp[2].p, 1.0, cw[0].p, cw[0].p, cw[1].p,
                                 cw[1].p, cw[2].p, cw[2].p)))
        return rc;
    DevBuf part;
    if ((rc = part.alloc(3 * 512))) return rc;
    for (int d = 0; d < 3; d++)
        hipLaunchKernelGGL(k_absmax, dim3(512), dim3(TPB), 0, st, g->face[d].p, g->ncell,
                           part.p + 512 * d);
    MPD_HIP(hipGetLastError());
    std::vector<double> h(3 * 512);
    MPD_HIP(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    MPD_HIP(hipStreamSynchronize(st));
    double mx = 0.0;
    for (double x : h) mx = std::fmax(mx, x);
    if (!(mx > 0.0)) return fail(MPD_ERR_ARG, "mpd_generate_wind: the wind is zero everywhere");
    const double f = w->cmax / mx;
    for (int d = 0; d < 3; d++)
        hipLaunchKernelGGL(k_scale, dim3(nblk(g->ncell)), dim3(TPB), 0, st, g->face[d].p,
                           g->ncell, f);
    MPD_HIP(hipGetLastError());
    MPD_HIP(hipStreamSynchronize(st));
    g->has_winds = true;
    return MPD_OK;
    MPD_API_END
}

extern "C" int mpd_generate_blobs(mpd_grid g, int target, int nblobs, const double* centres,
                                  const double* sigma, const double* amp, int periodic)
{
    MPD_API_BEGIN
    clear_error();
    if (!g || !centres || !sigma || !amp || nblobs < 1 || nblobs > MPD_MAX_BLOBS)
        return fail(MPD_ERR_ARG, "mpd_generate_blobs: bad arguments (1..%d blobs)", MPD_MAX_BLOBS);
    if (target != MPD_TARGET_FIELD && target != MPD_TARGET_EMISSION)
        return fail(MPD_ERR_ARG, "mpd_generate_blobs: target must be field or emission");
    if (target == MPD_TARGET_EMISSION && g->boundary != MPD_PERIODIC)
        return fail(MPD_ERR_UNSUPPORTED, "emission needs MPD_PERIODIC");
    BlobDev b{};
    b.nb = nblobs;
    b.periodic = periodic ? 1 : 0;
    for (int q = 0; q < nblobs; q++) {
        if (!(sigma[q] > 0)) return fail(MPD_ERR_ARG, "mpd_generate_blobs: sigma must be > 0");
        for (int e = 0; e < 3; e++) b.c[q][e] = centres[3 * q + e];
        b.inv2s2[q] = 0.5 / (sigma[q] * sigma[q]);
        b.amp[q] = amp[q];
    }
    int rc;
    double* out;
    if (target == MPD_TARGET_EMISSION) {
        if ((rc = g->q.alloc(g->ncell))) return rc;
        out = g->q.p;
    } else {
        out = g->c[g->cur].p;
    }
    hipLaunchKernelGGL(k_gen_blobs, dim3(nblk(g->ncell)), dim3(TPB), 0, stream(), g->nx, g->ny,
                       g->nz, b, out);
    MPD_HIP(hipGetLastError());
    MPD_HIP(hipStreamSynchronize(stream()));
    if (target == MPD_TARGET_EMISSION)
        g->has_q = true;
    else
        g->has_field = true;
    return MPD_OK;
    MPD_API_END
}
