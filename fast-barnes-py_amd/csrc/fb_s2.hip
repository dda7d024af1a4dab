// fb_s2.hip -- Barnes interpolation on the sphere S^2 (fastbarnes/interpolationS2.py).
//
//   optimized_convolution_S2: Lambert conformal projection of the points
//     (util/lambert_conformal.py:115-125), the optimized convolution on the fixed Lambert
//     grid (the batch plan of fb_api.hip), then the bilinear resampling of the Lambert
//     field to the lon/lat grid (interpolationS2.py:221-267)
//   naive_S2: O(N*W*H) Barnes with the spherical distance (interpolationS2.py:272-320)
// All arithmetic FP64; the device's sin/cos/tan/pow/acos/exp may differ from the host's
// libm in the last ulp, so results agree with the CPU restatement to rounding.
#include "fb_launch.h"

namespace fb {

namespace {

constexpr int kThreads = 256;
constexpr double kRad = 3.141592653589793 / 180.0;

struct LambertDev {
    double center_lon, n, n_inv, F, rho0;
};

__device__ __forceinline__ void lambert_map(const LambertDev &p, double lon, double lat, double &mx, double &my)
{
    const double rho = p.F / pow(tan((90.0 + lat) * (kRad / 2.0)), p.n);
    const double arg = p.n * (lon - p.center_lon) * kRad;
    mx = rho * sin(arg) / kRad;
    my = (p.rho0 - rho * cos(arg)) / kRad;
}

// to_map of the sample points: pts rows (lon, lat) -> Lambert map coordinates
__global__ __launch_bounds__(kThreads) void k_lambert_points(const double *__restrict__ pts,
                                                             double *__restrict__ out, int64_t n, LambertDev p)
{
    const int64_t k = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (k >= n) return;
    double mx, my;
    lambert_map(p, pts[2 * k], pts[2 * k + 1], mx, my);
    out[2 * k] = mx;
    out[2 * k + 1] = my;
}

// _resample: bilinear interpolation of the Lambert field at the projected lon/lat nodes.
// Nodes whose 2x2 Lambert neighbourhood leaves the Lambert grid are NaN (the reference
// does not check the bounds).
__global__ __launch_bounds__(kThreads) void k_resample(const float *__restrict__ lam, int64_t lw, int64_t lh,
                                                       double lx0, double ly0, double x0, double y0, double sx,
                                                       double sy, int64_t W, int64_t H, LambertDev p,
                                                       float *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (g >= W * H) return;
    const int64_t j = g / W, i = g - j * W;
    const double geoy = (double)j * sy + y0;
    const double geox = x0 + (double)i * sx;
    double mx, my;
    lambert_map(p, geox, geoy, mx, my);
    mx = (mx - lx0) / sx;
    my = (my - ly0) / sy;
    const int32_t ix = (int32_t)mx, iy = (int32_t)my;
    mx -= (double)ix;
    my -= (double)iy;
    if (!(ix >= 0 && iy >= 0 && ix + 1 < lw && iy + 1 < lh)) {
        out[g] = __builtin_nanf("");
        return;
    }
    const float *r0 = lam + (int64_t)iy * lw + ix;
    const float *r1 = r0 + lw;
    out[g] = (float)((1.0 - my) * (1.0 - mx) * (double)r0[0] + my * (1.0 - mx) * (double)r1[0] +
                     my * mx * (double)r1[1] + (1.0 - my) * mx * (double)r0[1]);
}

// naive_S2: one thread per node, point tiles (lon, sin lat, cos lat, value) staged in LDS
__global__ __launch_bounds__(kThreads) void k_naive_s2(const double *__restrict__ pts,
                                                       const double *__restrict__ val,
                                                       const double *__restrict__ offset, int64_t n,
                                                       double scale, double x0, double y0, double sx, double sy,
                                                       int64_t W, int64_t H, double *__restrict__ out)
{
    __shared__ double s_lon[kThreads], s_sin[kThreads], s_cos[kThreads], s_val[kThreads];
    const int tid = threadIdx.x;
    const int64_t g = (int64_t)blockIdx.x * kThreads + tid;
    const int64_t j = g / W, i = g - j * W;
    const double yc = y0 + (double)j * sy;
    const double xc = x0 + (double)i * sx;
    const double lat0 = yc * kRad;
    const double s0 = sin(lat0), c0 = cos(lat0);
    const double off = offset[0];
    double ws = 0.0, wt = 0.0;
    for (int64_t b = 0; b < n; b += kThreads) {
        __syncthreads();
        const int64_t k = b + tid;
        if (k < n) {
            const double lat1 = pts[2 * k + 1] * kRad;
            s_lon[tid] = pts[2 * k];
            s_sin[tid] = sin(lat1);
            s_cos[tid] = cos(lat1);
            s_val[tid] = val[k] - off;
        }
        __syncthreads();
        const int cnt = (int)min((int64_t)kThreads, n - b);
        if (g < W * H) {
            for (int t = 0; t < cnt; ++t) {
                double arg = s0 * s_sin[t] + c0 * s_cos[t] * cos((s_lon[t] - xc) * kRad);
                if (arg > 1.0) arg = 1.0;
                const double d = acos(arg) / kRad;
                const double w = exp(-d * d / scale);
                ws += w * s_val[t];
                wt += w;
            }
        }
    }
    if (g < W * H) out[g] = wt > 0.0 ? ws / wt + off : __builtin_nan("");
}

inline unsigned blocks(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

LambertDev to_dev(const double *proj) { return LambertDev{proj[0], proj[1], proj[2], proj[3], proj[4]}; }

}  // namespace

int launch_lambert_points(const double *d_pts, double *d_out, int64_t n, const double *proj, hipStream_t s)
{
    if (n <= 0) return FB_OK;
    hipLaunchKernelGGL(k_lambert_points, dim3(blocks(n)), dim3(kThreads), 0, s, d_pts, d_out, n, to_dev(proj));
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_resample(const float *d_lam, const int64_t *lam_size, const double *lam_x0, const double *x0,
                    const double *step, const int64_t *size, const double *proj, float *d_out, hipStream_t s)
{
    const int64_t cells = size[0] * size[1];
    if (cells <= 0) return FB_OK;
    hipLaunchKernelGGL(k_resample, dim3(blocks(cells)), dim3(kThreads), 0, s, d_lam, lam_size[0], lam_size[1],
                       lam_x0[0], lam_x0[1], x0[0], x0[1], step[0], step[1], size[0], size[1], to_dev(proj), d_out);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_naive_s2(const double *d_pts, const double *d_val, const double *d_offset, int64_t n, double sigma,
                    const double *x0, const double *step, const int64_t *size, double *d_out, hipStream_t s)
{
    const int64_t cells = size[0] * size[1];
    if (cells <= 0) return FB_OK;
    hipLaunchKernelGGL(k_naive_s2, dim3(blocks(cells)), dim3(kThreads), 0, s, d_pts, d_val, d_offset, n,
                       2 * (sigma * sigma), x0[0], x0[1], step[0], step[1], size[0], size[1], d_out);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

}  // namespace fb
